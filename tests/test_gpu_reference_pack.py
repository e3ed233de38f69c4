"""Reference-written packs through the HIP interpolation, against the
reference's own known answer.

The reference's exporter (``src/export_gpu_case.py:221-230``) writes a pack
directory together with ``dense_input.bin`` (FP16 [n_cols][16]) and
``reference.bin`` = FP16 coefficients x dense_input per channel, in float32
([output_size][16]); its binary accepts a kernel at max_abs <= 2e-2 and
rel_l2 <= 2e-3 (``nufft_spmma_bench.cu:1089-1101``).  Here the directory is
loaded unchanged by ``tsn_direction_load`` and run by ``tsn_interp`` (hi panel
= dense_input, residual panel = 0), gated at rel_l2 <= 2e-5 (and the
reference's own bounds).

The export is one real component.  Two complex directions are formed from it:
* real = imag = the reference directory: y_c = (A + iA)(x_re + i x_im), so
  Re y_c = ref[:, c] - ref[:, 8 + c] and Im y_c = ref[:, 8 + c] + ref[:, c];
* imag = a copy with zero coefficients (same structure): y_c = A x_re + i A x_im,
  so Re y_c = ref[:, c] and Im y_c = ref[:, 8 + c], reference.bin itself.
"""

import numpy as np
import pytest

from gpu_util import dev, need_gpu
from refpack import CASES, write_reference_dir

torch = pytest.importorskip("torch")
pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module", autouse=True)
def _gpu():
    need_gpu()


def _run(direction, dense_input, dense):
    from trajsparsenufft_amd.device import interp
    n_cols, rows = direction.cols, direction.rows
    hi = dev(dense_input.reshape(1, n_cols, 16))
    res = torch.zeros_like(hi)
    out = torch.zeros((8, rows), dtype=torch.complex64, device="cuda")
    interp(direction, hi, res, out, coils=8, dense=dense)
    torch.cuda.synchronize()
    return out.cpu().numpy()


def _gate(got, want):
    diff = np.abs(got - want)
    rel = float(np.linalg.norm(diff) / np.linalg.norm(want))
    assert np.isfinite(got).all()
    assert rel <= 2e-5, f"rel_l2 {rel:.3e}"
    assert float(diff.max()) <= 2e-2 and rel <= 2e-3  # the reference binary's own gate
    return rel


@pytest.mark.parametrize("dense", [False, True], ids=["smfmac", "dense_mfma"])
@pytest.mark.parametrize("case", CASES)
def test_reference_pack_reproduces_reference_bin(layout_golden, case, dense, tmp_path):
    from trajsparsenufft_amd.device import Direction
    g = layout_golden
    ref_dir = write_reference_dir(g, case, tmp_path / "ref")
    ref = g[f"{case}_reference"].reshape(-1, 16).astype(np.float64)
    x = g[f"{case}_dense_input"]
    print(f"[refpack] case={case} dense={dense} rows={ref.shape[0]}", flush=True)

    # imag = the same reference directory
    d_same = Direction.load(ref_dir, ref_dir)
    got = _run(d_same, x, dense)
    want = (ref[:, :8] - ref[:, 8:]).T + 1j * (ref[:, 8:] + ref[:, :8]).T
    _gate(got, want)

    # imag = zero coefficients on the reference's structure: reference.bin itself
    zero_dir = write_reference_dir(g, case, tmp_path / "zero")
    comp = g[f"{case}_tile_a_comp"]
    (zero_dir / "tile_a_comp.bin").write_bytes(np.zeros_like(comp).tobytes())
    d_real = Direction.load(ref_dir, zero_dir)
    got = _run(d_real, x, dense)
    want = ref[:, :8].T + 1j * ref[:, 8:].T
    _gate(got, want)


PRODUCTION_FILES = ("header", "group_offsets", "group_row_ids", "tile_col_ids", "tile_a_comp", "tile_pair_meta",
                    "dense_input", "reference")


@pytest.mark.parametrize("dense", [False, True], ids=["smfmac", "dense_mfma"])
@pytest.mark.parametrize("view", ["G", "G_T"])
def test_production_export_reproduces_its_reference_bin(view, dense, tmp_path):
    """The reference's production exporter (``src/export_torchkbnufft_gpu_case.py``
    ``export_component`` :377-489) writes a real and an imaginary directory,
    each with dense_input.bin and its own reference.bin (that component's FP16
    coefficients x dense_input).  Loaded as one complex direction:
    Re y_c = ref_re[:, c] - ref_im[:, 8 + c], Im y_c = ref_re[:, 8 + c] + ref_im[:, c]."""

    import os
    from trajsparsenufft_amd.device import Direction
    path = os.path.join(os.path.dirname(__file__), "golden", "production_golden.npz")
    with np.load(path, allow_pickle=False) as data:
        g = {k: data[k] for k in data.files}
    dirs = {}
    for comp in ("real", "imag"):
        d = tmp_path / comp
        d.mkdir()
        for name in PRODUCTION_FILES:
            (d / f"{name}.bin").write_bytes(g[f"{view}_morton_{comp}_{name}"].tobytes())
        dirs[comp] = d
    print(f"[refpack] production view={view} dense={dense}", flush=True)
    direction = Direction.load(dirs["real"], dirs["imag"])
    x = np.frombuffer(g[f"{view}_morton_real_dense_input"].tobytes(), dtype=np.float16)
    x_im = np.frombuffer(g[f"{view}_morton_imag_dense_input"].tobytes(), dtype=np.float16)
    assert np.array_equal(x.view(np.uint16), x_im.view(np.uint16))
    ref_re = np.frombuffer(g[f"{view}_morton_real_reference"].tobytes(), dtype=np.float32).reshape(-1, 16)
    ref_im = np.frombuffer(g[f"{view}_morton_imag_reference"].tobytes(), dtype=np.float32).reshape(-1, 16)
    ref_re, ref_im = ref_re.astype(np.float64), ref_im.astype(np.float64)
    got = _run(direction, x, dense)
    want = (ref_re[:, :8] - ref_im[:, 8:]).T + 1j * (ref_re[:, 8:] + ref_im[:, :8]).T
    _gate(got, want)
