"""Pack layout oracle: grouping, 2:4 arrangement, split24, compression.

TEST INFRASTRUCTURE ONLY (see ``oracle/__init__.py``).

Restates, function by function, the reference's structural CPU oracle and
exporters so the native packer (``trajsparsenufft_amd/csrc/packer.cpp``)
can be checked bit for bit:

* ``src/nufft_layout_oracle.py``: ``morton2d`` (:110-121),
  ``generate_trajectory`` (:124-149), ``build_interpolation`` (:159-206),
  ``transpose_sparse`` (:209-231), ``apply_sparse`` (:234-240), ``row_order``
  (:243-309), ``_column_masks`` (:312-320), ``_assign_columns_to_tiles``
  (:323-372), ``_arrange_24`` (:375-442), ``pack_view`` (:445-497),
  ``apply_packed`` (:500-513), checksums (:516-552), ``validate_24`` (:573-578).
* ``src/export_gpu_case.py``: ``pair_metadata`` (:41-48), ``compress_tile``
  (:51-74), ``split_illegal_tile`` (:77-107), ``split_legal_tiles_and_residual``
  (:110-155), header (:242-256).
* ``src/export_torchkbnufft_gpu_case.py``: ``optimized_assign_columns``
  (:102-162), ``optimized_arrange_24`` (:165-214), ``optimized_split24_pack``
  (:270-299, serial), ``compress_tiles_vectorized`` (:302-324),
  ``sparse_rows_from_coalesced`` (:327-353), ``materialize_component_tiles``
  (:356-384), per-component export (:387-436).
* ``experiments/packing/prepare_dense_control.py``: ``decode_row_metadata``
  (:27-36), dense expansion (:39-72).

The planners consume Python's ``random.Random`` stream in exactly the
reference's order, since the tile layout depends on it.

Parity: the functions from the first, second and fourth files are pinned
bit-exactly against golden fixtures written from the reference's own
importable (numpy-only) modules (``tests/golden/make_layout_golden.py``).  The
``optimized_*`` planner functions live in a reference module that imports
torchkbnufft at module level and cannot be imported here; they are pinned by
the structural invariants the reference enforces (lossless checksum, every
tile legal 2:4, no residual, identical real/imag structure) and by the
dense-expansion round trip — their exact tile order is "parity unpinned".
"""

from __future__ import annotations

import hashlib
import math
import random
import struct
import sys
from dataclasses import dataclass
from typing import Iterable, Sequence

import numpy as np

M_TILE = 16
K_TILE = 32
K_COMP = K_TILE // 2
MAGIC = 0x4E55464654323431  # "NUFFT241"
VERSION = 2
N_CHANNELS = 16
HEADER_FORMAT = "<Q11I"


@dataclass(frozen=True)
class SparseRows:
    """Row-sparse operator: ``cols[i]``/``vals[i]`` of local row i whose logical
    output row is ``row_ids[i]``; ``coords`` drive Morton grouping."""

    name: str
    output_size: int
    n_cols: int
    row_ids: np.ndarray
    coords: np.ndarray
    cols: tuple
    vals: tuple

    @property
    def n_active_rows(self) -> int:
        return len(self.cols)

    @property
    def nnz(self) -> int:
        return int(sum(len(c) for c in self.cols))


@dataclass
class Tile:
    row_ids: np.ndarray  # [16] int32, -1 = padding
    col_ids: np.ndarray  # [32] int32, -1 = padding
    values: np.ndarray   # [16, 32] float64
    legal_24: bool


# ----------------------------------------------------------------- geometry

def morton2d(x: int, y: int) -> int:
    def spread(v: int) -> int:
        v &= 0xFFFF
        v = (v | (v << 8)) & 0x00FF00FF
        v = (v | (v << 4)) & 0x0F0F0F0F
        v = (v | (v << 2)) & 0x33333333
        v = (v | (v << 1)) & 0x55555555
        return v
    return spread(x) | (spread(y) << 1)


def generate_trajectory(kind: str, grid_size: int, n_samples: int) -> np.ndarray:
    """Structural-oracle trajectories in grid units (reference :124-149)."""

    centre = (grid_size - 1) / 2.0
    rmax = 0.43 * grid_size
    if kind == "radial":
        spokes = max(16, int(round(math.sqrt(n_samples))))
        per_spoke = int(math.ceil(n_samples / spokes))
        step = math.pi * (3.0 - math.sqrt(5.0))
        radii = np.linspace(-rmax, rmax, per_spoke, endpoint=True)
        pts = []
        for s in range(spokes):
            c, sn = math.cos(s * step), math.sin(s * step)
            pts.extend((centre + r * c, centre + r * sn) for r in radii)
        return np.asarray(pts[:n_samples], dtype=np.float64)
    if kind == "spiral":
        t = (np.arange(n_samples, dtype=np.float64) + 0.5) / n_samples
        turns = max(6.0, grid_size / 8.0)
        radius = rmax * np.sqrt(t)
        angle = 2.0 * math.pi * turns * t
        return np.column_stack((centre + radius * np.cos(angle), centre + radius * np.sin(angle)))
    raise ValueError(f"unsupported trajectory kind: {kind}")


def _kb_weight(distance: float, width: int, beta: float) -> float:
    scaled = 2.0 * distance / width
    if abs(scaled) >= 1.0:
        return 0.0
    return float(np.i0(beta * math.sqrt(max(0.0, 1.0 - scaled * scaled))) / np.i0(beta))


def build_interpolation(trajectory: np.ndarray, grid_size: int, width: int,
                        oversampling: float = 2.0, periodic: bool = False) -> SparseRows:
    """Normalised separable KB matrix of the structural oracle (reference :159-206)."""

    beta = math.pi * math.sqrt(max((width / oversampling * (oversampling - 0.5)) ** 2 - 0.8, 0.01))
    half = width // 2
    cols_out, vals_out = [], []
    for xc, yc in trajectory:
        x0 = int(math.floor(xc)) - half + 1
        y0 = int(math.floor(yc)) - half + 1
        entries = []
        for gy in range(y0, y0 + width):
            if not periodic and not 0 <= gy < grid_size:
                continue
            wy = _kb_weight(yc - gy, width, beta)
            for gx in range(x0, x0 + width):
                if not periodic and not 0 <= gx < grid_size:
                    continue
                w = _kb_weight(xc - gx, width, beta) * wy
                if w != 0.0:
                    my = gy % grid_size if periodic else gy
                    mx = gx % grid_size if periodic else gx
                    entries.append((my * grid_size + mx, w))
        total = sum(w for _, w in entries)
        if total == 0.0:
            raise RuntimeError("empty interpolation row")
        entries.sort(key=lambda e: e[0])
        cols_out.append(np.asarray([c for c, _ in entries], dtype=np.int32))
        vals_out.append(np.asarray([w / total for _, w in entries], dtype=np.float64))
    return SparseRows("G", len(trajectory), grid_size * grid_size,
                      np.arange(len(trajectory), dtype=np.int32),
                      np.asarray(trajectory, dtype=np.float64), tuple(cols_out), tuple(vals_out))


def transpose_sparse(view: SparseRows, grid_size: int) -> SparseRows:
    """Exact active-row transpose; rows are active grid points ascending."""

    rows_of = [[] for _ in range(view.n_cols)]
    vals_of = [[] for _ in range(view.n_cols)]
    for local, (cols, vals) in enumerate(zip(view.cols, view.vals)):
        logical = int(view.row_ids[local])
        for c, v in zip(cols, vals):
            rows_of[int(c)].append(logical)
            vals_of[int(c)].append(float(v))
    active = [i for i, r in enumerate(rows_of) if r]
    coords = np.asarray([(i % grid_size, i // grid_size) for i in active], dtype=np.float64)
    return SparseRows("G_T", view.n_cols, view.output_size, np.asarray(active, dtype=np.int32),
                      coords, tuple(np.asarray(rows_of[i], dtype=np.int32) for i in active),
                      tuple(np.asarray(vals_of[i], dtype=np.float64) for i in active))


def sparse_rows_from_ell(name: str, cols: np.ndarray, vals: np.ndarray, n_cols: int,
                         coords: np.ndarray) -> SparseRows:
    """Rows with a fixed number of entries (the coalesced KB matrix) -> SparseRows
    (reference ``sparse_rows_from_coalesced`` :327-353 for row-sorted input)."""

    m = cols.shape[0]
    return SparseRows(name, m, n_cols, np.arange(m, dtype=np.int32),
                      np.asarray(coords, dtype=np.float64),
                      tuple(cols[i].astype(np.int32) for i in range(m)),
                      tuple(np.asarray(vals[i], dtype=np.float64) for i in range(m)))


def apply_sparse(view: SparseRows, dense: np.ndarray) -> np.ndarray:
    out = np.zeros((view.output_size, dense.shape[1]), dtype=np.complex128)
    for local, (cols, vals) in enumerate(zip(view.cols, view.vals)):
        out[int(view.row_ids[local])] = vals @ dense[cols]
    return out


# ----------------------------------------------------------------- planning

def row_order(view: SparseRows, policy: str, k_tile: int = K_TILE) -> list[int]:
    """Row grouping policies (reference :243-309)."""

    if policy == "trajectory":
        return list(range(view.n_active_rows))
    keys = [morton2d(max(0, int(round(view.coords[r, 0]))), max(0, int(round(view.coords[r, 1]))))
            for r in range(view.n_active_rows)]
    morton = sorted(range(view.n_active_rows), key=lambda r: keys[r])
    if policy == "morton":
        return morton
    if policy not in ("overlap", "packing256"):
        raise ValueError(f"unknown row grouping policy: {policy}")
    supports = [set(map(int, c)) for c in view.cols]
    window = 4 * M_TILE if policy == "overlap" else 16 * M_TILE
    ordered: list[int] = []
    for start in range(0, len(morton), window):
        pool = list(morton[start:start + window])
        while pool:
            group = [pool.pop(0)]
            union = set(supports[group[0]])
            while pool and len(group) < M_TILE:
                if policy == "overlap":
                    pick = max(range(len(pool)), key=lambda i: (
                        len(supports[pool[i]] & union), -len(supports[pool[i]] - union), -i))
                else:
                    nnz_now = sum(len(supports[r]) for r in group)
                    deg_now = max(len(supports[r]) for r in group)

                    def score(i: int):
                        cand = supports[pool[i]]
                        merged = union | cand
                        deg = max(deg_now, len(cand))
                        tiles = max(math.ceil(len(merged) / k_tile), math.ceil(deg / (k_tile // 2)), 1)
                        slots = tiles * M_TILE * k_tile // 2
                        return ((nnz_now + len(cand)) / slots, len(cand & union), -len(merged), -i)

                    pick = max(range(len(pool)), key=score)
                chosen = pool.pop(pick)
                group.append(chosen)
                union.update(supports[chosen])
            ordered.extend(group)
    return ordered


def column_masks(view: SparseRows, local_rows: Sequence[int]) -> dict:
    """Column -> bitmask of the group slots using it, in first-seen order."""

    masks: dict = {}
    for slot, r in enumerate(local_rows):
        if r < 0:
            continue
        for c in view.cols[r]:
            c = int(c)
            masks[c] = masks.get(c, 0) | (1 << slot)
    return masks


def _mask_bits(mask: int) -> np.ndarray:
    return np.asarray([(mask >> r) & 1 for r in range(M_TILE)])


def assign_columns_to_tiles(masks: dict, k_tile: int, seed: int) -> list[list[int]]:
    """Structural-oracle column partition (reference ``_assign_columns_to_tiles``)."""

    columns = list(masks)
    if not columns:
        return [[]]
    max_degree = max(sum(1 for m in masks.values() if m & (1 << r)) for r in range(M_TILE))
    t_min = max(math.ceil(len(columns) / k_tile), math.ceil(max_degree / (k_tile // 2)), 1)
    rng = random.Random(seed)
    for n_tiles in range(t_min, t_min + 9):
        for _ in range(24):
            tie = {c: rng.random() for c in columns}
            order = sorted(columns, key=lambda c: (-masks[c].bit_count(), tie[c]))
            tiles = [[] for _ in range(n_tiles)]
            counts = np.zeros((n_tiles, M_TILE), dtype=np.int16)
            ok = True
            for c in order:
                bits = _mask_bits(masks[c])
                cands = []
                for t in range(n_tiles):
                    if len(tiles[t]) >= k_tile:
                        continue
                    upd = counts[t] + bits
                    if int(upd.max(initial=0)) > k_tile // 2:
                        continue
                    dsq = float(np.sum(upd * upd - counts[t] ** 2))
                    cands.append((dsq + len(tiles[t]) / k_tile + rng.random() * 1e-6, t))
                if not cands:
                    ok = False
                    break
                _, t = min(cands)
                tiles[t].append(c)
                counts[t] += bits
            if ok:
                return [t for t in tiles if t]
    raise RuntimeError("failed to partition columns into half-density tiles")


def arrange_24(columns: Sequence[int], masks: dict, k_tile: int, seed: int) -> tuple[list[int], bool]:
    """Structural-oracle quartet arrangement (reference ``_arrange_24``)."""

    if len(columns) > k_tile:
        raise ValueError("too many columns for tile")
    n_groups = k_tile // 4
    rng = random.Random(seed)
    for _ in range(64):
        tie = {c: rng.random() for c in columns}
        order = sorted(columns, key=lambda c: (-masks[c].bit_count(), tie[c]))
        groups = [[] for _ in range(n_groups)]
        counts = np.zeros((n_groups, M_TILE), dtype=np.int8)
        ok = True
        for c in order:
            bits = np.asarray(_mask_bits(masks[c]), dtype=np.int8)
            choices = []
            for g in range(n_groups):
                if len(groups[g]) >= 4:
                    continue
                upd = counts[g] + bits
                if int(upd.max(initial=0)) > 2:
                    continue
                overlap = int(np.dot(counts[g], bits))
                choices.append((overlap * 100.0 + len(groups[g]) + rng.random() * 1e-6, g))
            if not choices:
                ok = False
                break
            _, g = min(choices)
            groups[g].append(c)
            counts[g] += bits
        if not ok:
            continue
        for g in groups:
            g.extend([-1] * (4 - len(g)))
        return [c for g in groups for c in g], True
    return list(columns) + [-1] * (k_tile - len(columns)), False


def optimized_assign_columns(masks: dict, seed: int) -> list[list[int]]:
    """Production column partition (reference :102-162): vectorised scoring and
    one random draw per tile per column, 8 restarts per tile count."""

    columns = list(masks)
    if not columns:
        return [[]]
    bits = {c: np.asarray(_mask_bits(m), dtype=np.int16) for c, m in masks.items()}
    max_degree = max(sum(int(b[r]) for b in bits.values()) for r in range(M_TILE))
    t_min = max(math.ceil(len(columns) / K_TILE), math.ceil(max_degree / (K_TILE // 2)), 1)
    rng = random.Random(seed)
    for n_tiles in range(t_min, t_min + 9):
        for _ in range(8):
            tie = {c: rng.random() for c in columns}
            order = sorted(columns, key=lambda c: (-masks[c].bit_count(), tie[c]))
            tiles = [[] for _ in range(n_tiles)]
            sizes = np.zeros(n_tiles, dtype=np.int16)
            counts = np.zeros((n_tiles, M_TILE), dtype=np.int16)
            ok = True
            for c in order:
                upd = counts + bits[c][None, :]
                valid = (sizes < K_TILE) & (np.max(upd, axis=1) <= K_TILE // 2)
                if not np.any(valid):
                    ok = False
                    break
                score = np.sum(upd * upd - counts * counts, axis=1, dtype=np.int64).astype(np.float64)
                score += sizes.astype(np.float64) / K_TILE
                score += np.asarray([rng.random() * 1e-6 for _ in range(n_tiles)])
                score[~valid] = np.inf
                t = int(np.argmin(score))
                tiles[t].append(c)
                sizes[t] += 1
                counts[t] = upd[t]
            if ok:
                return [t for t in tiles if t]
    raise RuntimeError("optimized column assignment failed")


def optimized_arrange_24(columns: list[int], masks: dict, seed: int) -> tuple[list[int], bool]:
    """Production quartet arrangement (reference :165-214): 16 restarts."""

    n_groups = K_TILE // 4
    bits = {c: np.asarray(_mask_bits(masks[c]), dtype=np.int8) for c in columns}
    rng = random.Random(seed)
    for _ in range(16):
        tie = {c: rng.random() for c in columns}
        order = sorted(columns, key=lambda c: (-masks[c].bit_count(), tie[c]))
        groups = [[] for _ in range(n_groups)]
        sizes = np.zeros(n_groups, dtype=np.int8)
        counts = np.zeros((n_groups, M_TILE), dtype=np.int8)
        ok = True
        for c in order:
            upd = counts + bits[c][None, :]
            valid = (sizes < 4) & (np.max(upd, axis=1) <= 2)
            if not np.any(valid):
                ok = False
                break
            score = (counts @ bits[c]).astype(np.float64) * 100.0
            score += sizes.astype(np.float64)
            score += np.asarray([rng.random() * 1e-6 for _ in range(n_groups)])
            score[~valid] = np.inf
            g = int(np.argmin(score))
            groups[g].append(c)
            sizes[g] += 1
            counts[g] = upd[g]
        if ok:
            for g in groups:
                g.extend([-1] * (4 - len(g)))
            return [c for g in groups for c in g], True
    return list(columns) + [-1] * (K_TILE - len(columns)), False


def validate_24(values: np.ndarray) -> tuple[bool, int]:
    """(legal, violation count) of a [16, K] value block."""

    viol = 0
    for off in range(0, values.shape[1], 4):
        cnt = np.count_nonzero(values[:, off:off + 4], axis=1)
        viol += int(np.maximum(cnt - 2, 0).sum())
    return viol == 0, viol


def _group_rows(view: SparseRows, local_rows: Sequence[int]):
    row_ids = np.asarray([int(view.row_ids[r]) if r >= 0 else -1 for r in local_rows], dtype=np.int32)
    maps = [dict(zip(map(int, view.cols[r]), map(float, view.vals[r]))) if r >= 0 else {}
            for r in local_rows]
    return row_ids, maps


def _tile_from(row_ids, maps, arranged) -> np.ndarray:
    values = np.zeros((M_TILE, len(arranged)), dtype=np.float64)
    for slot, m in enumerate(maps):
        for j, c in enumerate(arranged):
            if c >= 0:
                values[slot, j] = m.get(c, 0.0)
    return values


def _group_slices(order: list[int]):
    for gi, off in enumerate(range(0, len(order), M_TILE)):
        rows = list(order[off:off + M_TILE])
        yield gi, rows + [-1] * (M_TILE - len(rows))


def pack_view(view: SparseRows, policy: str, k_tile: int, seed: int) -> list[Tile]:
    """Structural-oracle packer (reference ``pack_view`` :445-497)."""

    tiles: list[Tile] = []
    for gi, rows in _group_slices(row_order(view, policy, k_tile)):
        masks = column_masks(view, rows)
        row_ids, maps = _group_rows(view, rows)
        for ti, cols in enumerate(assign_columns_to_tiles(masks, k_tile, seed + gi * 1009)):
            arranged, legal = arrange_24(cols, masks, k_tile, seed + gi * 1009 + ti * 9176)
            tiles.append(Tile(row_ids.copy(), np.asarray(arranged, dtype=np.int32),
                              _tile_from(row_ids, maps, arranged), legal))
    return tiles


def split_illegal_tile(tile: Tile) -> tuple[Tile, Tile]:
    """split24: first two nonzeros of every row quartet -> layer 0, rest -> layer 1."""

    l0 = np.zeros_like(tile.values)
    l1 = np.zeros_like(tile.values)
    for r in range(M_TILE):
        for off in range(0, tile.values.shape[1], 4):
            pos = np.flatnonzero(tile.values[r, off:off + 4])
            for p in pos[:2]:
                l0[r, off + p] = tile.values[r, off + p]
            for p in pos[2:]:
                l1[r, off + p] = tile.values[r, off + p]
    a = Tile(tile.row_ids.copy(), tile.col_ids.copy(), l0, True)
    b = Tile(tile.row_ids.copy(), tile.col_ids.copy(), l1, True)
    if not (validate_24(l0)[0] and validate_24(l1)[0]) or not np.array_equal(l0 + l1, tile.values):
        raise RuntimeError("split24 failed")
    return a, b


def optimized_split24_pack(view: SparseRows, policy: str, seed: int) -> list[Tile]:
    """Production planner (reference ``optimized_split24_pack`` / ``_optimized_pack_group``)."""

    tiles: list[Tile] = []
    for gi, rows in _group_slices(row_order(view, policy, K_TILE)):
        row_ids, maps = _group_rows(view, rows)
        masks: dict = {}
        for slot, m in enumerate(maps):
            for c in m:
                masks[c] = masks.get(c, 0) | (1 << slot)
        for ti, cols in enumerate(optimized_assign_columns(masks, seed + gi * 1009)):
            arranged, legal = optimized_arrange_24(cols, masks, seed + gi * 1009 + ti * 9176)
            tile = Tile(row_ids.copy(), np.asarray(arranged, dtype=np.int32),
                        _tile_from(row_ids, maps, arranged), legal)
            tiles.extend([tile] if legal else split_illegal_tile(tile))
    return tiles


def split_legal_tiles_and_residual(view: SparseRows, tiles: list[Tile], illegal_mode: str = "split24"):
    """Group offsets, group row ids, legal tiles and CSR residual (reference :110-155)."""

    offsets = [0]
    rows = []
    legal: list[Tile] = []
    res_cols = [[] for _ in range(view.output_size)]
    res_vals = [[] for _ in range(view.output_size)]
    previous = None
    for tile in tiles:
        key = tuple(map(int, tile.row_ids))
        if key != previous:
            if previous is not None:
                offsets.append(len(legal))
            rows.append(tile.row_ids.astype(np.int32, copy=True))
            previous = key
        if validate_24(tile.values)[0]:
            legal.append(tile)
        elif illegal_mode == "split24":
            legal.extend(split_illegal_tile(tile))
        else:
            for slot, lr in enumerate(tile.row_ids):
                if lr < 0:
                    continue
                for j, lc in enumerate(tile.col_ids):
                    v = float(tile.values[slot, j])
                    if lc >= 0 and v != 0.0:
                        res_cols[int(lr)].append(int(lc))
                        res_vals[int(lr)].append(v)
    offsets.append(len(legal))
    res_ptr = np.zeros(view.output_size + 1, dtype=np.uint32)
    res_ptr[1:] = np.cumsum([len(c) for c in res_cols], dtype=np.uint64)
    return (np.asarray(offsets, dtype=np.uint32), np.stack(rows).astype(np.int32), legal, res_ptr,
            np.asarray([c for r in res_cols for c in r], dtype=np.int32),
            np.asarray([v for r in res_vals for v in r], dtype=np.float16))


def materialize_component_tiles(structure: list[Tile], component: SparseRows) -> list[Tile]:
    """Fill real or imag values into the shared structure (reference :356-384)."""

    lookup = {int(r): i for i, r in enumerate(component.row_ids)}
    maps = [dict(zip(map(int, c), map(float, v))) for c, v in zip(component.cols, component.vals)]
    out = []
    for st in structure:
        values = np.zeros_like(st.values)
        for slot, lr in enumerate(st.row_ids):
            if lr < 0:
                continue
            m = maps[lookup[int(lr)]]
            for j, lc in enumerate(st.col_ids):
                if lc >= 0 and st.values[slot, j] != 0.0:
                    values[slot, j] = m.get(int(lc), 0.0)
        if not validate_24(values)[0]:
            raise RuntimeError("materialized component is not legal 2:4")
        out.append(Tile(st.row_ids.copy(), st.col_ids.copy(), values, True))
    return out


# ----------------------------------------------------------------- compression

def pair_metadata(row_meta: np.ndarray) -> np.ndarray:
    """Row words -> the pair-packed u32[16] of one tile (reference :41-48)."""

    out = np.zeros(M_TILE, dtype=np.uint32)
    for g in range(8):
        a, b = int(row_meta[g]), int(row_meta[g + 8])
        out[2 * g] = (a & 0xFFFF) | ((b & 0xFFFF) << 16)
        out[2 * g + 1] = ((a >> 16) & 0xFFFF) | (b & 0xFFFF0000)
    return out


def compress_tile(values: np.ndarray) -> tuple[np.ndarray, np.ndarray]:
    """One [16, 32] tile -> (f16[16, 16] kept values, u32[16] pair metadata)."""

    rounded = values.astype(np.float16)
    comp = np.zeros((M_TILE, K_COMP), dtype=np.float16)
    row_meta = np.zeros(M_TILE, dtype=np.uint32)
    for r in range(M_TILE):
        word = 0
        for ch in range(K_TILE // 4):
            four = rounded[r, 4 * ch:4 * ch + 4]
            keep = [int(p) for p in np.flatnonzero(four)]
            if len(keep) > 2:
                raise ValueError("tile is not legal 2:4 after FP16 rounding")
            for p in range(4):
                if len(keep) == 2:
                    break
                if p not in keep:
                    keep.append(p)
            p0, p1 = sorted(keep)
            comp[r, 2 * ch] = four[p0]
            comp[r, 2 * ch + 1] = four[p1]
            word |= (p0 | (p1 << 2)) << (4 * ch)
        row_meta[r] = word
    return comp, pair_metadata(row_meta)


def bf16_bits(values: np.ndarray) -> np.ndarray:
    """uint16 bfloat16 bits of float32(values), round to nearest even (NaN kept
    quiet).  Not a reference function: the library's BF16 tile variant."""

    u = np.asarray(values, dtype=np.float32).view(np.uint32).astype(np.uint64)
    nan = (u & 0x7FFFFFFF) > 0x7F800000
    r = ((u + 0x7FFF + ((u >> 16) & 1)) >> 16) & 0xFFFF
    r = np.where(nan, (u >> 16) | 0x40, r)
    return r.astype(np.uint16)


def bf16_decode(bits: np.ndarray) -> np.ndarray:
    return (np.asarray(bits, dtype=np.uint16).astype(np.uint32) << np.uint32(16)).view(np.float32)


def compress_tiles_vectorized(values: np.ndarray, fmt: str = "fp16") -> tuple[np.ndarray, np.ndarray]:
    """[T, 16, 32] float64 -> (f16[T,16,16], u32[T,16]) (reference :302-324).

    fmt "bf16" applies the same selection to BF16-rounded values and returns
    the kept values as uint16 BF16 bits."""

    t = values.shape[0]
    if fmt == "bf16":
        quads = bf16_bits(values.astype(np.float32)).reshape(t, M_TILE, K_TILE // 4, 4)
        nonzero = (quads & 0x7FFF) != 0  # -0 counts as zero, as for FP16
    else:
        quads = values.astype(np.float16).reshape(t, M_TILE, K_TILE // 4, 4)
        nonzero = quads != 0
    pos = np.arange(4, dtype=np.uint32)
    priority = np.where(nonzero, pos, pos + 4)
    sel = np.sort(np.argsort(priority, axis=-1)[..., :2], axis=-1).astype(np.uint32)
    comp = np.take_along_axis(quads, sel, axis=-1).reshape(t, M_TILE, K_COMP)
    codes = sel[..., 0] | (sel[..., 1] << np.uint32(2))
    shifts = (np.arange(K_TILE // 4, dtype=np.uint32) * np.uint32(4))[None, None, :]
    row_meta = np.bitwise_or.reduce(codes << shifts, axis=2)
    pm = np.zeros((t, M_TILE), dtype=np.uint32)
    for g in range(8):
        a, b = row_meta[:, g], row_meta[:, g + 8]
        pm[:, 2 * g] = (a & 0xFFFF) | ((b & 0xFFFF) << 16)
        pm[:, 2 * g + 1] = (a >> 16) | (b & 0xFFFF0000)
    return comp, pm


def decode_row_metadata(pair_meta: np.ndarray) -> np.ndarray:
    """u32[T,16] pair metadata -> u32[T,16] per-row 4-bit chunk codes."""

    out = np.zeros((pair_meta.shape[0], M_TILE), dtype=np.uint32)
    for g in range(8):
        a, b = pair_meta[:, 2 * g], pair_meta[:, 2 * g + 1]
        out[:, g] = (a & 0xFFFF) | ((b & 0xFFFF) << 16)
        out[:, g + 8] = (a >> 16) | (b & 0xFFFF0000)
    return out


def expand_tiles(comp: np.ndarray, pair_meta: np.ndarray) -> np.ndarray:
    """(f16[T,16,16], u32[T,16]) -> dense f16[T,16,32] (the dense-control tiles)."""

    t = comp.shape[0]
    rm = decode_row_metadata(pair_meta)
    dense = np.zeros((t, M_TILE, K_TILE), dtype=comp.dtype)
    ti = np.arange(t)[:, None]
    ri = np.arange(M_TILE)[None, :]
    for ch in range(K_TILE // 4):
        code = (rm >> np.uint32(4 * ch)) & np.uint32(0xF)
        p0 = (code & 3).astype(np.int64)
        p1 = ((code >> 2) & 3).astype(np.int64)
        dense[ti, ri, 4 * ch + p0] = comp[:, :, 2 * ch]
        dense[ti, ri, 4 * ch + p1] = comp[:, :, 2 * ch + 1]
    return dense


# ----------------------------------------------------------------- checksums

def _hash_records(records: Iterable) -> str:
    h = hashlib.sha256()
    for row, col, val in sorted(records):
        h.update(struct.pack("<qqd", row, col, val))
    return h.hexdigest()


def sparse_checksum(view: SparseRows) -> str:
    return _hash_records((int(view.row_ids[i]), int(c), float(v))
                         for i, (cs, vs) in enumerate(zip(view.cols, view.vals)) for c, v in zip(cs, vs))


def packed_checksum(tiles: list[Tile]) -> str:
    return _hash_records((int(t.row_ids[r]), int(t.col_ids[j]), float(t.values[r, j]))
                         for t in tiles for r in range(M_TILE) if t.row_ids[r] >= 0
                         for j in range(t.values.shape[1]) if t.col_ids[j] >= 0 and t.values[r, j] != 0.0)


def canonical_operator_checksum(view: SparseRows, transposed: bool) -> str:
    if not transposed:
        return sparse_checksum(view)
    return _hash_records((int(c), int(view.row_ids[i]), float(v))
                         for i, (cs, vs) in enumerate(zip(view.cols, view.vals)) for c, v in zip(cs, vs))


def apply_packed(n_out: int, tiles: list[Tile], dense: np.ndarray) -> np.ndarray:
    out = np.zeros((n_out, dense.shape[1]), dtype=np.complex128)
    for t in tiles:
        valid = np.flatnonzero(t.col_ids >= 0)
        if not len(valid):
            continue
        part = t.values[:, valid] @ dense[t.col_ids[valid]]
        for slot, lr in enumerate(t.row_ids):
            if lr >= 0:
                out[int(lr)] += part[slot]
    return out


# ----------------------------------------------------------------- export

def header_bytes(output_size: int, n_cols: int, active_rows: int, nnz: int, groups: int,
                 tiles: int, residual_nnz: int = 0) -> bytes:
    return struct.pack(HEADER_FORMAT, MAGIC, VERSION, N_CHANNELS, output_size, n_cols, active_rows,
                       nnz, groups, tiles, residual_nnz, M_TILE, K_TILE)


def export_arrays(structure_view: SparseRows, structure_tiles: list[Tile], offsets, row_ids,
                  component: SparseRows, fmt: str = "fp16") -> dict:
    """Arrays of one production component directory (reference :387-436)."""

    tiles = materialize_component_tiles(structure_tiles, component)
    col_ids = np.stack([t.col_ids for t in tiles]).astype(np.int32)
    col_ids[col_ids < 0] = 0
    comp, meta = compress_tiles_vectorized(np.stack([t.values for t in tiles]), fmt)
    return {
        "header.bin": header_bytes(component.output_size, component.n_cols, component.n_active_rows,
                                   component.nnz, len(offsets) - 1, len(tiles)),
        "group_offsets.bin": np.ascontiguousarray(offsets, dtype=np.uint32),
        "group_row_ids.bin": np.ascontiguousarray(row_ids, dtype=np.int32),
        "tile_col_ids.bin": col_ids,
        "tile_a_comp.bin": comp if fmt == "bf16" else comp.astype(np.float16),
        "tile_pair_meta.bin": meta.astype(np.uint32),
    }


def pack_components(real: SparseRows, imag: SparseRows, policy: str = "morton",
                    seed: int = 20260823, fmt: str = "fp16") -> dict:
    """Full production pack of one direction: {"real": arrays, "imag": arrays}.

    The structure view carries ones in place of the values, as the reference
    packs ``TKBN_STRUCTURE_G`` (reference :561-596)."""

    structure = SparseRows(real.name, real.output_size, real.n_cols, real.row_ids, real.coords,
                           real.cols, tuple(np.ones(len(c)) for c in real.cols))
    tiles = optimized_split24_pack(structure, policy, seed)
    offsets, row_ids, legal, _, res_cols, _ = split_legal_tiles_and_residual(structure, tiles, "split24")
    if len(res_cols):
        raise RuntimeError("split24 structure export left residual entries")
    return {"real": export_arrays(structure, legal, offsets, row_ids, real, fmt),
            "imag": export_arrays(structure, legal, offsets, row_ids, imag, fmt)}


if __name__ == "__main__":  # pragma: no cover
    sys.exit("oracle module; import it from tests")
