"""bench.py's command line on the CPU: it measures one GPU and says so when
asked for more (no GPU work starts)."""

import subprocess
import sys
from pathlib import Path

ROOT = Path(__file__).resolve().parents[1]


def test_more_than_one_gpu_is_refused_with_a_message():
    proc = subprocess.run([sys.executable, str(ROOT / "bench.py"), "--gpus", "2"], capture_output=True, text=True,
                          timeout=60, cwd=ROOT)
    assert proc.returncode == 2
    assert proc.stdout == ""
    assert "one GPU only" in proc.stderr and "--gpus 2" in proc.stderr


def test_help_describes_the_workload():
    proc = subprocess.run([sys.executable, str(ROOT / "bench.py"), "--help"], capture_output=True, text=True,
                          timeout=60, cwd=ROOT)
    assert proc.returncode == 0
    assert "256x256 image on a 512x512 grid" in proc.stdout and "--steps" in proc.stdout
