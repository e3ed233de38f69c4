"""Pack directories written by the reference, rebuilt from the committed fixtures.

``tests/golden/layout_golden.npz`` holds, for two cases, every file the
reference's exporter wrote (``src/export_gpu_case.py:242-280``): the 52-byte
``header.bin`` and the ``group_offsets``, ``group_row_ids``,
``tile_col_ids``, ``tile_a_comp``, ``tile_pair_meta`` and ``residual_*``
arrays, plus its known-answer pair ``dense_input.bin`` / ``reference.bin``
(``:221-230``: reference = FP16 coefficients x dense_input per channel, in
float32).  ``write_reference_dir`` lays them out again as the directory the
reference's reader takes (``nufft_integrated_bench.cu:135-226``), byte for
byte as written.
"""

from __future__ import annotations

from pathlib import Path

import numpy as np

CASES = ("export_split24", "export_csr_residual")
FILES = ("group_offsets", "group_row_ids", "tile_col_ids", "tile_a_comp", "tile_pair_meta",
         "residual_row_ptr", "residual_cols", "residual_vals", "dense_input", "reference")


def write_reference_dir(golden: dict, case: str, path: Path) -> Path:
    path.mkdir(parents=True, exist_ok=True)
    (path / "header.bin").write_bytes(golden[f"{case}_header"].tobytes())
    for name in FILES:
        (path / f"{name}.bin").write_bytes(np.ascontiguousarray(golden[f"{case}_{name}"]).tobytes())
    return path


def header_fields(golden: dict, case: str) -> dict:
    raw = golden[f"{case}_header"].tobytes()
    q = np.frombuffer(raw[:8], dtype="<u8")[0]
    i = np.frombuffer(raw[8:], dtype="<u4")
    names = ("version", "n_channels", "output_size", "n_cols", "active_rows", "nnz", "groups", "tiles",
             "residual_nnz", "m_tile", "k_tile")
    out = {"magic": int(q)}
    out.update({n: int(v) for n, v in zip(names, i)})
    return out
