"""Paired-timing harness and quality admission on the CPU: statistics pinned by
closed forms, order alternation and aggregation with stand-in commands (small
Python programs printing the reference CLI's JSON line), row admission."""

import json
import math
import sys

import numpy as np
import pytest

from trajsparsenufft_amd import paired


def test_quantile_linear_interpolation():
    assert paired.quantile([3.0, 1.0, 2.0], 0.5) == 2.0
    assert paired.quantile([0.0, 10.0], 0.25) == 2.5
    assert paired.quantile([5.0], 0.975) == 5.0
    # numpy's default ("linear") quantile is the same rule
    v = list(np.random.default_rng(0).random(17))
    for p in (0.025, 0.5, 0.975):
        assert paired.quantile(v, p) == pytest.approx(float(np.quantile(v, p)), rel=1e-15)


def test_bootstrap_geomean_closed_forms():
    const = paired.bootstrap_geomean_ci([1.5] * 6, seed=1, draws=2000)
    assert const["geomean"] == pytest.approx(1.5) and const["ci95_low"] == pytest.approx(1.5)
    assert const["ci95_high"] == pytest.approx(1.5)
    two = paired.bootstrap_geomean_ci([1.0, 4.0], seed=2, draws=4000)
    assert two["geomean"] == pytest.approx(2.0)
    # a resample of two values takes geomeans 1, 2 or 4 only; the interval spans them
    assert two["ci95_low"] == pytest.approx(1.0) and two["ci95_high"] == pytest.approx(4.0)
    a = paired.bootstrap_geomean_ci([1.1, 1.3, 1.2, 1.25], seed=3, draws=3000)
    b = paired.bootstrap_geomean_ci([1.1, 1.3, 1.2, 1.25], seed=3, draws=3000)
    assert a == b and a["ci95_low"] <= a["geomean"] <= a["ci95_high"]
    with pytest.raises(ValueError):
        paired.bootstrap_geomean_ci([1.0, 0.0], seed=1)


def test_parse_result_and_showpids():
    out = "log line\n{\"variant\":\"x\",\"correct\":false,\"us\":1}\n{\"correct\":true,\"us\":12.5}\n"
    assert paired.parse_result(out)["us"] == 12.5
    with pytest.raises(RuntimeError):
        paired.parse_result("{\"correct\":true}\n")
    text = ("============ KFD Processes ============\nKFD process information:\n"
            "PID\tPROCESS NAME\tGPU(s)\tVRAM USED\tSDMA USED\tCU OCCUPANCY\n"
            "4242\tpython3\t1\t1048576\t0\t0\n977\ttsn_cg\t1\t2048\t0\t0\n")
    assert paired.parse_showpids(text, [4242]) == ["977\ttsn_cg\t1\t2048\t0\t0"]
    assert paired.parse_showpids("No KFD PIDs currently running\n") == []


def _fake(us_by_case):
    """A command printing the CLI's JSON line with a fixed time per case."""

    def command(case):
        return [sys.executable, "-c", f"import json; print(json.dumps({{'correct': True, 'us': {us_by_case[case]}}}))"]
    return command


def test_run_paired_alternates_and_aggregates(tmp_path):
    fast = _fake({"radial": 100.0, "spiral": 200.0})
    slow = _fake({"radial": 150.0, "spiral": 260.0})
    summary = paired.run_paired({"split16": fast, "fp32": slow}, ["radial", "spiral"], tmp_path / "c", pairs=3,
                                audit=False, variant_env={"fp32": {"TSN_PRECISION": "fp32"}})
    rows = [json.loads(l) for l in (tmp_path / "c" / "paired_rows.jsonl").read_text().splitlines()]
    assert len(rows) == 6
    for r in rows:  # order alternates with (pair + case index)
        first = "split16" if (r["pair"] + ["radial", "spiral"].index(r["case"])) % 2 == 0 else "fp32"
        assert r["order"][0] == first
    # per-pair ratio = sum(B) / sum(A) = 410 / 300
    assert summary["geomean_ratio"] == pytest.approx(410.0 / 300.0)
    assert summary["ci95_low"] == pytest.approx(410.0 / 300.0) and summary["all_wins"]
    assert summary["decision"] == "ACCEPT"  # 1.367 >= 1.20 in every pair
    cases = {c["case"]: c for c in summary["case_summaries"]}
    assert cases["radial"]["geomean_ratio"] == pytest.approx(1.5)
    assert cases["spiral"]["geomean_ratio"] == pytest.approx(1.3)
    records = json.loads((tmp_path / "c" / "records.json").read_text())
    assert len(records) == 12 and all(r["env"] == ({"TSN_PRECISION": "fp32"} if r["variant"] == "fp32" else {})
                                      for r in records)
    log = (tmp_path / "c" / "raw_processes" / "pair0_radial_fp32.log").read_text()
    assert log.startswith("$ TSN_PRECISION=fp32 ")
    with pytest.raises(FileExistsError):
        paired.run_paired({"a": fast, "b": slow}, ["radial"], tmp_path / "c", pairs=1, audit=False)


def test_run_paired_stops_on_a_failed_process(tmp_path):
    ok = _fake({"radial": 100.0})

    def bad(case):
        return [sys.executable, "-c", "import sys; sys.exit(3)"]
    with pytest.raises(RuntimeError, match="exit 3"):
        paired.run_paired({"a": ok, "b": bad}, ["radial"], tmp_path / "d", pairs=1, audit=False)


def test_psnr_and_ssim_properties():
    rng = np.random.default_rng(4)
    img = rng.random((48, 40))
    assert paired.ssim(img, img, 1.0) == pytest.approx(1.0)
    assert paired.psnr(img, img, 1.0) == math.inf
    # PSNR closed form: constant offset d gives MSE d^2
    assert paired.psnr(img, img + 0.01, 1.0) == pytest.approx(40.0)
    # a constant offset leaves every variance and covariance unchanged: SSIM is the luminance term
    d = 0.05
    f = lambda a: __import__("scipy.ndimage", fromlist=["uniform_filter"]).uniform_filter(a, 7, mode="reflect")  # noqa: E731
    mu = f(img)
    c1 = 0.01 ** 2
    lum = ((2 * mu * (mu + d) + c1) / (mu * mu + (mu + d) ** 2 + c1))[3:-3, 3:-3].mean()
    assert paired.ssim(img, img + d, 1.0) == pytest.approx(float(lum), rel=1e-9)
    noise = rng.standard_normal(img.shape)
    s1, s3 = paired.ssim(img, img + 0.1 * noise, 1.0), paired.ssim(img, img + 0.3 * noise, 1.0)
    assert s3 < s1 < 0.99


def _row(case, frame, **over):
    x = np.exp(1j * np.linspace(0, 3, 64 * 64)).reshape(64, 64) * np.linspace(1, 2, 64)[None, :]
    row = paired.quality_row(x * (1 + 1e-7), x, None, eager_graph_image_bitwise=True,
                             eager_graph_residuals_bitwise=True, case=case, frame=frame, cg_iterations=10)
    row.update(over)
    return row


def test_validate_rows_admission():
    rows = [_row(c, f) for c in ("radial", "spiral") for f in range(2)]
    assert all(r["application_vs_fp32_pass"] for r in rows)
    paired.validate_rows(rows, ["radial", "spiral"], 2)
    with pytest.raises(RuntimeError, match="missing"):
        paired.validate_rows(rows[:-1], ["radial", "spiral"], 2)
    with pytest.raises(RuntimeError, match="missing"):
        paired.validate_rows(rows + [rows[0]], ["radial", "spiral"], 2)
    with pytest.raises(RuntimeError, match="determinism"):
        paired.validate_rows(rows[:-1] + [_row("spiral", 1, eager_graph_reconstruction_bitwise=False)],
                             ["radial", "spiral"], 2)
    with pytest.raises(RuntimeError, match="CG steps"):
        paired.validate_rows(rows[:-1] + [_row("spiral", 1, cg_iterations=5)], ["radial", "spiral"], 2)
    worse = paired.quality_row(np.ones((16, 16)) * 2, np.ones((16, 16)) + np.eye(16), None,
                               eager_graph_image_bitwise=True, eager_graph_residuals_bitwise=True)
    assert not worse["application_vs_fp32_pass"]
