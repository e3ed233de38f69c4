"""Runtime-input directories: the reference file set, manifest and checks."""

import json

import numpy as np
import pytest

from trajsparsenufft_amd import runtime


def _case(coils=3, samples=50, n=8, seed=0):
    rng = np.random.default_rng(seed)
    meas = (rng.standard_normal((coils, samples)) + 1j * rng.standard_normal((coils, samples))).astype(np.complex64)
    maps = (rng.standard_normal((coils, n, n)) + 1j * rng.standard_normal((coils, n, n))).astype(np.complex64)
    return meas, maps, rng.random((n, n)).astype(np.float32)


def test_round_trip_and_manifest(tmp_path):
    meas, maps, scaling = _case()
    truth = (maps[0] * 2).astype(np.complex64)
    man = runtime.write_runtime_dir(tmp_path / "c", meas, maps, scaling, truth=truth, name="radial", frame=3)
    assert man["coils"] == 3 and man["samples"] == 50 and man["image_size"] == [8, 8] and man["grid_size"] == [16, 16]
    names = [e["file"] for e in man["entries"]]
    assert set(runtime.REQUIRED) <= set(names) and "truth.c64.bin" in names
    got = runtime.read_runtime_dir(tmp_path / "c")
    np.testing.assert_array_equal(got["measurement"], meas)
    np.testing.assert_array_equal(got["sensitivity_maps"], maps)
    np.testing.assert_array_equal(got["scaling"], scaling)
    np.testing.assert_array_equal(got["density"], np.ones(50, np.float32))
    # raw little-endian payloads, as the CLI and the reference binary read them
    np.testing.assert_array_equal(np.fromfile(tmp_path / "c" / "measurement.c64.bin", np.complex64), meas.ravel())
    # a corrupted payload fails its manifest
    with open(tmp_path / "c" / "scaling.f32.bin", "r+b") as f:
        f.write(b"\x00\x01")
    with pytest.raises(ValueError, match="fails its manifest"):
        runtime.read_runtime_dir(tmp_path / "c")
    with pytest.raises(ValueError):
        runtime.write_runtime_dir(tmp_path / "bad", meas, maps[:2], scaling)


def test_root_manifest(tmp_path):
    meas, maps, scaling = _case()
    for frame in (0, 1):
        fdir = tmp_path / "ocmr" / f"frame{frame:02d}"
        cases = [runtime.write_runtime_dir(fdir / name, meas, maps, scaling, name=name, frame=frame)
                 for name in ("spiral65", "radial65")]
        (fdir / "manifest.json").write_text(json.dumps(cases))
    with pytest.raises(RuntimeError, match="expected 3"):
        runtime.build_root_manifest(tmp_path, 3)
    rec = runtime.build_root_manifest(tmp_path, 2, cases=("spiral65", "radial65"))
    assert rec["schema_version"] == 1 and len(rec["frame_manifests"]) == 2
    with pytest.raises(FileExistsError):
        runtime.build_root_manifest(tmp_path, 2)
