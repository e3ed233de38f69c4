"""Paired timing campaign and quality admission.

Two reconstruction commands (variants) are timed as separate processes,
pair by pair and case by case, with the order inside each (pair, case)
alternating so that neither variant always runs on a warmer GPU.  Before each
process the GPU is audited for processes that are not this campaign's, and
the board state is logged.  The per-pair ratio is the sum of variant B's
times over the sum of variant A's; the campaign reports the geometric mean
of those ratios with a 95% percentile-bootstrap interval, per case as well.
A campaign only runs once a quality matrix has been admitted: every row
finite, eager and Graph reconstructions bitwise equal, and the image gates
(SSIM >= 0.9999, PSNR >= 60 dB against the FP32 solve, truth NRMSE within
0.002 of it).

Reference: ``experiments/cufinufft_baseline/run_paired_fullcg.py:45-67``
(quantile, bootstrap, result parsing), ``:162-250`` (order alternation,
foreign-process audit, raw logs, paired rows), ``:258-330`` (aggregation and
the ACCEPT rule), and ``tools/quality_contract.py:10-37`` (row admission).
The commands are any programs printing the reference CLI's JSON result line
(``{"correct": true, "us": ...}``), such as ``tsn_cg`` under two
``TSN_PRECISION`` values.

SSIM follows Wang et al. (2004) with the defaults of scikit-image's
``structural_similarity`` (7x7 uniform window, sample covariance, K1 0.01,
K2 0.03, mean over the window-valid interior); scikit-image is not
installed here, so that equality is not checked against it.
"""

from __future__ import annotations

import argparse
import hashlib
import itertools
import json
import math
import os
import random
import shutil
import statistics
import subprocess
import time
from pathlib import Path
from typing import Callable, Sequence

import numpy as np

BOOTSTRAP_DRAWS = 50000


def sha256(path: Path) -> str:
    digest = hashlib.sha256()
    with Path(path).open("rb") as stream:
        for chunk in iter(lambda: stream.read(1 << 20), b""):
            digest.update(chunk)
    return digest.hexdigest()


def quantile(values: Sequence[float], probability: float) -> float:
    """Sample quantile by linear interpolation between order statistics
    (numpy's default rule: position probability * (n - 1))."""

    return float(np.quantile(np.asarray(values, dtype=np.float64), probability))


def bootstrap_geomean_ci(ratios: Sequence[float], seed: int, draws: int = BOOTSTRAP_DRAWS) -> dict:
    """Geometric mean of positive ratios and its 95% percentile-bootstrap interval."""

    if not ratios or any(not (r > 0.0) or not math.isfinite(r) for r in ratios):
        raise ValueError("ratios must be positive and finite")
    logs = [math.log(r) for r in ratios]
    rng = random.Random(seed)
    n = len(logs)
    means = [math.exp(statistics.fmean([logs[rng.randrange(n)] for _ in range(n)])) for _ in range(draws)]
    return {"geomean": math.exp(statistics.fmean(logs)), "ci95_low": quantile(means, 0.025),
            "ci95_high": quantile(means, 0.975)}


def parse_result(stdout: str) -> dict:
    """The result object of a run: the first line of output that parses as a
    JSON object with ``correct`` true and a time ``us``."""

    candidates = (json.loads(line) for line in stdout.splitlines() if line.lstrip().startswith("{"))
    found = next((obj for obj in candidates if obj.get("correct") is True and "us" in obj), None)
    if found is None:
        raise RuntimeError("no correct timed result line in the output (missing)")
    return found


# ------------------------------------------------------------ GPU audit

def _rocm_smi() -> str | None:
    return shutil.which("rocm-smi") or (os.path.exists("/opt/rocm/bin/rocm-smi") and "/opt/rocm/bin/rocm-smi") or None


def foreign_processes(allowed_pids: Sequence[int] = ()) -> list[str]:
    """Lines of ``rocm-smi --showpids`` naming GPU processes not in allowed_pids.

    rocm-smi lists compute processes as ``PID  NAME  GPU(s)  VRAM ...`` rows;
    a row whose first field is a number other than an allowed PID is foreign."""

    tool = _rocm_smi()
    if tool is None:
        raise RuntimeError("rocm-smi is required for the exclusive-GPU audit")
    text = subprocess.run([tool, "--showpids"], text=True, capture_output=True, check=True).stdout
    return parse_showpids(text, allowed_pids)


def parse_showpids(text: str, allowed_pids: Sequence[int] = ()) -> list[str]:
    allowed = {int(p) for p in allowed_pids}
    foreign = []
    for line in text.splitlines():
        fields = line.split()
        if fields and fields[0].isdigit() and int(fields[0]) not in allowed:
            foreign.append(line.strip())
    return foreign


def gpu_state() -> str:
    """Board snapshot: use, memory, temperature, power and clocks (rocm-smi)."""

    tool = _rocm_smi()
    if tool is None:
        return "rocm-smi unavailable"
    proc = subprocess.run([tool, "--showuse", "--showmemuse", "--showtemp", "--showpower", "--showclocks", "--csv"],
                          text=True, capture_output=True, check=False)
    return proc.stdout.strip() or proc.stderr.strip()


# ------------------------------------------------------------ campaign

Command = Callable[[str], Sequence[str]]


def run_paired(variants: dict[str, Command], cases: Sequence[str], output: Path, pairs: int = 6,
               audit: bool = True, accept_ci_low: float = 1.20, seed: int = 20260825,
               variant_env: dict[str, dict] | None = None) -> dict:
    """Time the two variants (A first in ``variants``, B second) over ``pairs``
    x ``cases``; ratio = B time / A time (B slower than A gives ratios > 1).
    ``variant_env`` adds environment variables to a variant's processes.

    Writes ``raw_processes/*.log``, ``paired_rows.jsonl``, ``records.json`` and
    ``summary.json`` under ``output`` (which must be empty or absent)."""

    if len(variants) != 2:
        raise ValueError("a paired campaign compares exactly two variants")
    if pairs <= 0 or not cases:
        raise ValueError("need at least one pair and one case")
    name_a, name_b = list(variants)
    output = Path(output)
    if output.exists() and any(output.iterdir()):
        raise FileExistsError(f"output directory {output} already holds files; give an empty or new one")
    raw = output / "raw_processes"
    raw.mkdir(parents=True, exist_ok=True)
    records, snapshots, paired = [], [], []
    for pair in range(pairs):
        for case_index, case in enumerate(cases):
            order = (name_a, name_b) if (pair + case_index) % 2 == 0 else (name_b, name_a)
            payloads = {}
            for variant in order:
                if audit:
                    foreign = foreign_processes([os.getpid()])
                    snapshots.append({"pair": pair, "case": case, "variant": variant, "foreign": foreign,
                                      "gpu_state": gpu_state()})
                    if foreign:
                        raise RuntimeError(f"foreign GPU process before {pair}/{case}/{variant}: {foreign}")
                command = [str(c) for c in variants[variant](case)]
                extra = (variant_env or {}).get(variant, {})
                started = time.time_ns()
                proc = subprocess.run(command, text=True, capture_output=True, check=False,
                                      env={**os.environ, **extra} if extra else None)
                ended = time.time_ns()
                log = raw / f"pair{pair}_{case}_{variant}.log"
                shown = " ".join(f"{k}={v}" for k, v in extra.items()) + (" " if extra else "") + " ".join(command)
                log.write_text("$ " + shown + "\n\n[stdout]\n" + proc.stdout + "\n[stderr]\n" + proc.stderr)
                if proc.returncode != 0:
                    raise RuntimeError(f"{variant} failed at {pair}/{case}: exit {proc.returncode} (see {log})")
                payload = parse_result(proc.stdout)
                payloads[variant] = payload
                records.append({"pair": pair, "case": case, "variant": variant, "order": list(order),
                                "command": command, "env": extra, "started_ns": started, "ended_ns": ended, "payload": payload,
                                "raw_log": str(log)})
            row = {"pair": pair, "case": case, "order": list(order), f"{name_a}_us": float(payloads[name_a]["us"]),
                   f"{name_b}_us": float(payloads[name_b]["us"])}
            row["ratio"] = row[f"{name_b}_us"] / row[f"{name_a}_us"]
            paired.append(row)
            with (output / "paired_rows.jsonl").open("a") as stream:
                stream.write(json.dumps(row, sort_keys=True) + "\n")
    return summarize(paired, name_a, name_b, cases, pairs, output, records, snapshots, accept_ci_low, seed)


def summarize(paired: list[dict], name_a: str, name_b: str, cases: Sequence[str], pairs: int, output: Path,
              records: list, snapshots: list, accept_ci_low: float, seed: int) -> dict:
    aggregate = []
    for pair in range(pairs):
        rows = [r for r in paired if r["pair"] == pair]
        sum_a = sum(r[f"{name_a}_us"] for r in rows)
        sum_b = sum(r[f"{name_b}_us"] for r in rows)
        aggregate.append({"pair": pair, f"{name_a}_sum_us": sum_a, f"{name_b}_sum_us": sum_b,
                          "ratio": sum_b / sum_a})
    ratios = [r["ratio"] for r in aggregate]
    interval = bootstrap_geomean_ci(ratios, seed)
    per_case = []
    for i, case in enumerate(cases):
        case_ratios = [r["ratio"] for r in paired if r["case"] == case]
        ci = bootstrap_geomean_ci(case_ratios, seed + 75 + i)
        per_case.append({"case": case, "geomean_ratio": ci["geomean"], "ci95_low": ci["ci95_low"],
                         "ci95_high": ci["ci95_high"], "wins": sum(v > 1.0 for v in case_ratios)})
    all_wins = all(v > 1.0 for v in ratios)
    summary = {"variants": [name_a, name_b], "ratio": f"{name_b}_us / {name_a}_us", "pairs": pairs,
               "cases": list(cases), "case_summaries": per_case, "aggregate_rows": aggregate,
               "geomean_ratio": interval["geomean"], "ci95_low": interval["ci95_low"],
               "ci95_high": interval["ci95_high"], "wins": sum(v > 1.0 for v in ratios), "all_wins": all_wins,
               "snapshots": snapshots,
               "decision": "ACCEPT" if all_wins and interval["ci95_low"] >= accept_ci_low else "MEASURED"}
    (output / "records.json").write_text(json.dumps(records, indent=2, sort_keys=True) + "\n")
    (output / "summary.json").write_text(json.dumps(summary, indent=2, sort_keys=True) + "\n")
    return summary


# ------------------------------------------------------------ quality admission

def psnr(reference: np.ndarray, candidate: np.ndarray, data_range: float) -> float:
    mse = float(np.mean((np.asarray(reference, np.float64) - np.asarray(candidate, np.float64)) ** 2))
    return math.inf if mse == 0.0 else 10.0 * math.log10(data_range * data_range / mse)


def ssim(reference: np.ndarray, candidate: np.ndarray, data_range: float, win_size: int = 7) -> float:
    """Mean structural similarity of two 2-D images (uniform window)."""

    from scipy.ndimage import uniform_filter
    x = np.asarray(reference, np.float64)
    y = np.asarray(candidate, np.float64)
    if x.shape != y.shape or x.ndim != 2 or min(x.shape) < win_size:
        raise ValueError("ssim needs two equal 2-D images at least one window wide")
    n = win_size * win_size
    cov_norm = n / (n - 1.0)
    f = lambda a: uniform_filter(a, size=win_size, mode="reflect")  # noqa: E731
    ux, uy = f(x), f(y)
    vx = cov_norm * (f(x * x) - ux * ux)
    vy = cov_norm * (f(y * y) - uy * uy)
    vxy = cov_norm * (f(x * y) - ux * uy)
    c1 = (0.01 * data_range) ** 2
    c2 = (0.03 * data_range) ** 2
    s = ((2 * ux * uy + c1) * (2 * vxy + c2)) / ((ux * ux + uy * uy + c1) * (vx + vy + c2))
    pad = (win_size - 1) // 2
    return float(s[pad:-pad, pad:-pad].mean())


def rel_l2(a, b) -> float:
    a = np.asarray(a)
    b = np.asarray(b)
    return float(np.linalg.norm(a - b) / max(float(np.linalg.norm(b)), 1e-30))


def quality_row(native: np.ndarray, fp32: np.ndarray, truth: np.ndarray | None, *, eager_graph_image_bitwise: bool,
                eager_graph_residuals_bitwise: bool, **fields) -> dict:
    """One admission row: the candidate's magnitude image against the FP32 solve
    (and the truth when given), with the gates evaluated."""

    nm, fm = np.abs(native), np.abs(fp32)
    data_range = max(float(fm.max() - fm.min()), 1e-12)
    row = dict(fields)
    row.update({
        "finite": bool(np.isfinite(native).all()),
        "native_vs_fp32_magnitude_nrmse": rel_l2(nm, fm),
        "native_vs_fp32_ssim": ssim(fm, nm, data_range),
        "native_vs_fp32_psnr_db": psnr(fm, nm, data_range),
        "eager_graph_reconstruction_bitwise": bool(eager_graph_image_bitwise),
        "eager_graph_residuals_bitwise": bool(eager_graph_residuals_bitwise),
    })
    if truth is not None:
        tm = np.abs(truth)
        row["native_truth_nrmse"] = rel_l2(nm, tm)
        row["fp32_truth_nrmse"] = rel_l2(fm, tm)
    row["application_vs_fp32_pass"] = bool(
        row["finite"] and row["native_vs_fp32_ssim"] >= 0.9999 and row["native_vs_fp32_psnr_db"] >= 60.0
        and (truth is None or row["native_truth_nrmse"] - row["fp32_truth_nrmse"] <= 0.002))
    return row


def validate_rows(rows: list[dict], cases: Sequence[str], frames: int, iterations: int = 10) -> None:
    """Admit a complete (case x frame) matrix or raise: no missing or duplicate
    rows, the given CG length, finite metrics, eager/Graph bitwise, gates met."""

    expected = set(itertools.product(cases, range(frames)))
    keys = [(r["case"], r["frame"]) for r in rows]
    if len(keys) != len(expected) or set(keys) != expected:
        raise RuntimeError("quality rows do not cover each (case, frame) exactly once: some are missing or repeated")
    for row in rows:
        if row["cg_iterations"] != iterations:
            raise RuntimeError(f"quality row does not use {iterations} CG steps")
        if not all(row[k] is True for k in ("finite", "application_vs_fp32_pass", "eager_graph_reconstruction_bitwise",
                                            "eager_graph_residuals_bitwise")):
            raise RuntimeError(f"row {row['case']}/{row['frame']} fails its image gates or eager/Graph determinism")
        metrics = [float(row[k]) for k in ("native_vs_fp32_magnitude_nrmse", "native_vs_fp32_ssim")]
        if not all(math.isfinite(v) for v in metrics):
            raise RuntimeError(f"row {row['case']}/{row['frame']} has a metric that is not finite")


# ------------------------------------------------------------ CLI

def tsn_cg_command(cli: Path, packed_root: Path, runtime_root: Path, scale, warmup: int, iters: int,
                   graph: bool = False) -> Command:
    """Command of one tsn_cg run per case (argument order of nufft_fp16x2_cg);
    ``scale`` is one operator scale or a mapping case -> scale."""

    def command(case: str) -> list[str]:
        pack = Path(packed_root) / case
        s = scale[case] if isinstance(scale, dict) else scale
        return [str(cli), str(pack / "G/real"), str(pack / "G/imag"), str(pack / "G_T/real"), str(pack / "G_T/imag"),
                str(Path(runtime_root) / case), f"{float(s):.9g}", str(warmup), str(iters), "0",
                "1" if graph else "0", "10"]
    return command


def main(argv=None) -> int:
    parser = argparse.ArgumentParser(description="paired timing of two tsn_cg precisions")
    parser.add_argument("--cli", type=Path, required=True, help="tsn_cg binary")
    parser.add_argument("--packed-root", type=Path, required=True, help="<case>/G/{real,imag}, <case>/G_T/{real,imag}")
    parser.add_argument("--runtime-root", type=Path, required=True, help="<case>/ runtime-input directories")
    parser.add_argument("--cases", required=True, help="comma-separated case names")
    parser.add_argument("--variants", default="split16,dense16", help="two TSN_PRECISION values: A,B")
    parser.add_argument("--operator-scale", type=float, required=True)
    parser.add_argument("--output", type=Path, required=True)
    parser.add_argument("--pairs", type=int, default=6)
    parser.add_argument("--warmup", type=int, default=5)
    parser.add_argument("--iters", type=int, default=20)
    parser.add_argument("--no-audit", action="store_true", help="skip the exclusive-GPU audit")
    args = parser.parse_args(argv)
    a, b = args.variants.split(",")
    command = tsn_cg_command(args.cli, args.packed_root, args.runtime_root, args.operator_scale, args.warmup,
                             args.iters)
    summary = run_paired({a: command, b: command}, args.cases.split(","), args.output, args.pairs,
                         audit=not args.no_audit, variant_env={a: {"TSN_PRECISION": a}, b: {"TSN_PRECISION": b}})
    print(json.dumps({k: summary[k] for k in ("variants", "ratio", "geomean_ratio", "ci95_low", "ci95_high",
                                              "wins", "decision")}))
    return 0


if __name__ == "__main__":
    raise SystemExit(main())
