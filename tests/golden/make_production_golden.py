#!/usr/bin/env python3
"""Write the production-planner fixtures from the reference's own exporter.

Run in the development container only (the reference tree is not present on
GPU machines):  python tests/golden/make_production_golden.py /root/reference

The reference's production pack path -- ``optimized_split24_pack`` /
``_optimized_pack_group`` / ``optimized_assign_columns`` /
``optimized_arrange_24`` (the planner), ``materialize_component_tiles``,
``compress_tiles_vectorized``, ``sparse_rows_from_coalesced`` and
``export_component`` (the writer) -- lives in
``src/export_torchkbnufft_gpu_case.py``.  Those functions use only numpy and
the standard library, but the module imports torchkbnufft (absent here) at
its top, so it cannot be imported whole.  This script parses that file and
compiles exactly those function definitions, in a namespace holding the names
they use from the standard library, numpy and the importable numpy-only
reference modules (``src/nufft_layout_oracle.py``, ``src/export_gpu_case.py``).
Nothing else of the module runs.

The coefficients fed to the reference exporter are this repository's own KB
matrix of a small golden-angle radial trajectory (the values torchkbnufft's
``calc_tensor_spmatrix`` would supply; the planner depends only on their
structure).  The fixture keeps the inputs (trajectory, per-row columns and
float32 real/imag values) and every file the reference exporter wrote for
each (view, policy, component): header, group offsets and row ids, tile
columns, compressed FP16 values, pair metadata, dense_input and reference.
"""

from __future__ import annotations

import __future__
import argparse
import ast
import importlib.util
import json
import math
import multiprocessing as mp
import random
import sys
import tempfile
import time
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parents[2]
WANTED = ("optimized_assign_columns", "optimized_arrange_24", "_optimized_pack_group", "optimized_split24_pack",
          "compress_tiles_vectorized", "sparse_rows_from_coalesced", "materialize_component_tiles",
          "export_component")
FILES = ("header.bin", "group_offsets.bin", "group_row_ids.bin", "tile_col_ids.bin", "tile_a_comp.bin",
         "tile_pair_meta.bin", "dense_input.bin", "reference.bin")
CASE = {"image": 24, "spokes": 12, "readout": 48, "numpoints": 6, "seed": 20260823}
RUNS = (("G", "morton"), ("G_T", "morton"), ("G_T", "overlap"))


def load(root: Path, relative: str, name: str):
    spec = importlib.util.spec_from_file_location(name, root / relative)
    module = importlib.util.module_from_spec(spec)
    sys.modules[name] = module
    spec.loader.exec_module(module)
    return module


def production_functions(root: Path) -> dict:
    src = root / "src" / "export_torchkbnufft_gpu_case.py"
    tree = ast.parse(src.read_text(), filename=str(src))
    keep = [node for node in tree.body
            if (isinstance(node, ast.FunctionDef) and node.name in WANTED)
            or (isinstance(node, ast.AnnAssign) and getattr(node.target, "id", "") == "_parallel_view")]
    found = {n.name for n in keep if isinstance(n, ast.FunctionDef)}
    missing = set(WANTED) - found
    if missing:
        raise RuntimeError(f"functions not found in {src}: {sorted(missing)}")
    oracle = load(root, "src/nufft_layout_oracle.py", "src.nufft_layout_oracle")
    export = load(root, "src/export_gpu_case.py", "src.export_gpu_case")
    ns = {"np": np, "math": math, "random": random, "mp": mp, "time": time, "json": json,
          "hashlib": __import__("hashlib"), "struct": __import__("struct"), "Path": Path}
    for name in ("K_TILE", "MAGIC", "N_CHANNELS", "VERSION", "compress_tile", "split_illegal_tile",
                 "split_legal_tiles_and_residual", "write_array"):
        ns[name] = getattr(export, name)
    for name in ("M_TILE", "PackedView", "PackedTile", "SparseRows", "pack_view", "row_order",
                 "transpose_sparse", "validate_24"):
        ns[name] = getattr(oracle, name)
    module = ast.Module(body=keep, type_ignores=[])
    code = compile(module, str(src), "exec", flags=__future__.annotations.compiler_flag, dont_inherit=True)
    exec(code, ns)
    return ns


def main() -> int:
    parser = argparse.ArgumentParser()
    parser.add_argument("reference", type=Path)
    parser.add_argument("--output", type=Path, default=Path(__file__).with_name("production_golden.npz"))
    args = parser.parse_args()
    root = args.reference.resolve()
    sys.path.insert(0, str(root))
    sys.path.insert(0, str(ROOT))
    ns = production_functions(root)

    from trajsparsenufft_amd import pack, synth
    image, k = CASE["image"], 2 * CASE["image"]
    traj = synth.radial(CASE["spokes"], CASE["readout"], True)
    kb = pack.KbMatrix.build(traj, (image, image), (k, k), numpoints=CASE["numpoints"])
    cols, re_, im_ = kb.arrays()
    rows, width = cols.shape
    out: dict[str, np.ndarray] = {"traj": traj, "kb_cols": cols, "kb_re": re_, "kb_im": im_,
                                  "case": np.frombuffer(json.dumps(CASE).encode(), dtype=np.uint8)}
    # the coalesced COO structure calc_tensor_spmatrix hands the exporter:
    # row-major, columns ascending within a row
    indices = np.stack([np.repeat(np.arange(rows), width), cols.reshape(-1)]).astype(np.int64)
    shape = (rows, k * k)
    coords = np.mod(traj.astype(np.float64) * k + k / 2.0, k)
    components = {
        "real": ns["sparse_rows_from_coalesced"]("TKBN_REAL_G", indices, re_.reshape(-1), shape, coords),
        "imag": ns["sparse_rows_from_coalesced"]("TKBN_IMAG_G", indices, im_.reshape(-1), shape, coords),
    }
    structure = ns["sparse_rows_from_coalesced"]("TKBN_STRUCTURE_G", indices,
                                                 np.ones(indices.shape[1], dtype=np.float32), shape, coords)
    with tempfile.TemporaryDirectory() as tmp:
        for view, policy in RUNS:
            if view == "G_T":
                sview = ns["transpose_sparse"](structure, k)
                comps = {c: ns["transpose_sparse"](v, k) for c, v in components.items()}
            else:
                sview, comps = structure, components
            packed = ns["optimized_split24_pack"](sview, policy, CASE["seed"], 1)
            offsets, row_ids, tiles, _, res_cols, _ = ns["split_legal_tiles_and_residual"](sview, packed, "split24")
            assert len(res_cols) == 0
            rng = np.random.default_rng(CASE["seed"])
            dense_input = np.clip(rng.standard_normal((sview.n_cols, ns["N_CHANNELS"])), -3.0, 3.0).astype(np.float16)
            for comp, cview in comps.items():
                d = Path(tmp) / f"{view}_{policy}_{comp}"
                manifest = ns["export_component"](d, comp, cview, offsets, row_ids, tiles, dense_input, {}, False)
                for name in FILES:
                    if policy != "morton" and name in ("dense_input.bin", "reference.bin"):
                        continue  # the known answer is kept for the production policy only
                    out[f"{view}_{policy}_{comp}_{name[:-4]}"] = np.frombuffer((d / name).read_bytes(), dtype=np.uint8)
                out[f"{view}_{policy}_{comp}_utilization"] = np.float64(manifest["fp16_useful_slot_utilization"])
            print(f"{view} {policy}: {len(tiles)} tiles, {len(offsets) - 1} groups")
    np.savez_compressed(args.output, **out)
    print(f"wrote {args.output}")
    return 0


if __name__ == "__main__":
    raise SystemExit(main())
