"""Runtime-input directories: the per-case files the reconstruction CLI reads.

Same file set, dtypes and manifest schema as the reference's exporter
(``experiments/data_prep/export_real_runtime_inputs.py:24-34, 97-127``), so a
directory prepared by either side feeds the other's ``nufft_fp16x2_cg`` /
``tsn_cg`` unchanged; and the root manifest of
``tools/build_runtime_manifest.py:21-57`` over per-frame manifests.

Files (all little-endian, C order):

* ``measurement.c64.bin``       complex64 [coils, samples], raw (unscaled) k-space
* ``sensitivity_maps.c64.bin``  complex64 [coils, N0, N1]
* ``scaling.f32.bin``           float32 [N0, N1], the real apodization
* ``density.f32.bin``           float32 [samples] (all ones: no density weighting)
* optional ``forward_coils.c64.bin``, ``truth.f32.bin``, ``truth.c64.bin``
"""

from __future__ import annotations

import hashlib
import json
from pathlib import Path

import numpy as np

REQUIRED = ("measurement.c64.bin", "sensitivity_maps.c64.bin", "scaling.f32.bin", "density.f32.bin")


def sha256(path: Path) -> str:
    digest = hashlib.sha256()
    with Path(path).open("rb") as stream:
        for chunk in iter(lambda: stream.read(1 << 20), b""):
            digest.update(chunk)
    return digest.hexdigest()


def _write(directory: Path, name: str, array: np.ndarray) -> dict:
    path = directory / name
    data = np.ascontiguousarray(array)
    data.tofile(path)
    return {"file": name, "shape": list(data.shape), "dtype": str(data.dtype), "bytes": path.stat().st_size,
            "sha256": sha256(path)}


def write_runtime_dir(directory, measurement: np.ndarray, smaps: np.ndarray, scaling: np.ndarray, *,
                      density: np.ndarray | None = None, forward_coils: np.ndarray | None = None,
                      truth: np.ndarray | None = None, name: str = "case", frame: int = 0,
                      grid_size=None, trajectory_file: str = "", extra: dict | None = None) -> dict:
    """Write one runtime-input directory plus its ``manifest.json``; returns the manifest."""

    directory = Path(directory)
    directory.mkdir(parents=True, exist_ok=True)
    meas = np.asarray(measurement, np.complex64)
    maps = np.asarray(smaps, np.complex64)
    if meas.ndim != 2 or maps.ndim != 3 or meas.shape[0] != maps.shape[0]:
        raise ValueError("measurement must be [coils, samples] and smaps [coils, N0, N1]")
    if np.asarray(scaling).shape != maps.shape[1:]:
        raise ValueError("scaling must be [N0, N1]")
    image = list(maps.shape[1:])
    entries = [
        _write(directory, "measurement.c64.bin", meas),
        _write(directory, "sensitivity_maps.c64.bin", maps),
        _write(directory, "scaling.f32.bin", np.asarray(scaling, np.float32)),
        _write(directory, "density.f32.bin",
               np.ones(meas.shape[1], np.float32) if density is None else np.asarray(density, np.float32)),
    ]
    if forward_coils is not None:
        entries.append(_write(directory, "forward_coils.c64.bin", np.asarray(forward_coils, np.complex64)))
    if truth is not None:
        entries.append(_write(directory, "truth.f32.bin", np.abs(truth).astype(np.float32)))
        entries.append(_write(directory, "truth.c64.bin", np.asarray(truth, np.complex64)))
    manifest = {
        "name": name, "frame": int(frame), "samples": int(meas.shape[1]), "coils": int(meas.shape[0]),
        "image_size": image, "grid_size": list(grid_size) if grid_size else [2 * n for n in image],
        "trajectory_file": trajectory_file, "measurement_scale": 1,
        "density_contract": "all ones" if density is None else "given", "entries": entries,
    }
    if extra:
        manifest.update(extra)
    (directory / "manifest.json").write_text(json.dumps(manifest, indent=2) + "\n", encoding="utf-8")
    return manifest


def read_runtime_dir(directory, verify: bool = True) -> dict:
    """Arrays of a runtime-input directory, shaped by its manifest; with
    ``verify`` every payload must match the manifest's size and SHA-256."""

    directory = Path(directory)
    manifest = json.loads((directory / "manifest.json").read_text(encoding="utf-8"))
    if manifest.get("measurement_scale", 1) != 1:
        raise ValueError("runtime measurements must be unscaled (measurement_scale 1)")
    names = {e["file"] for e in manifest["entries"]}
    missing = [f for f in REQUIRED if f not in names]
    if missing:
        raise ValueError(f"runtime directory lacks {missing}")
    out = {"manifest": manifest}
    for e in manifest["entries"]:
        path = directory / e["file"]
        if verify and (path.stat().st_size != e["bytes"] or sha256(path) != e["sha256"]):
            raise ValueError(f"runtime payload fails its manifest: {path}")
        out[e["file"].split(".")[0]] = np.fromfile(path, dtype=np.dtype(e["dtype"])).reshape(e["shape"])
    return out


def build_root_manifest(root, expected: int, cases=None) -> dict:
    """Root ``manifest.json`` over ``**/frame??/manifest.json`` (each a list of
    case manifests in case subdirectories), after checking every payload."""

    root = Path(root).resolve()
    output = root / "manifest.json"
    if output.exists():
        raise FileExistsError(f"refusing to overwrite {output}")
    manifests = sorted(root.glob("**/frame??/manifest.json"))
    if len(manifests) != expected:
        raise RuntimeError(f"found {len(manifests)} frame manifests under {root}; expected {expected}")
    for path in manifests:
        frame_cases = json.loads(path.read_text(encoding="utf-8"))
        if cases is not None and {c["name"] for c in frame_cases} != set(cases):
            raise RuntimeError(f"incomplete trajectory matrix in {path}")
        for case in frame_cases:
            if case["measurement_scale"] != 1:
                raise RuntimeError("runtime measurements must be unscaled")
            for e in case["entries"]:
                payload = path.parent / case["name"] / e["file"]
                if payload.stat().st_size != e["bytes"] or sha256(payload) != e["sha256"]:
                    raise RuntimeError(f"runtime payload fails its manifest: {payload}")
    record = {"schema_version": 1, "root": ".",
              "frame_manifests": [{"path": str(p.relative_to(root)), "sha256": sha256(p)} for p in manifests]}
    output.write_text(json.dumps(record, indent=2, sort_keys=True) + "\n", encoding="utf-8")
    return record


__all__ = ["write_runtime_dir", "read_runtime_dir", "build_root_manifest", "sha256", "REQUIRED"]
