"""Build libtsn.so (HIP for gfx950 + C++ host code) in-tree with hipcc.

Usage: python -m trajsparsenufft_amd.build [--jobs N] [--force]

Every HIP translation unit is compiled with
``-Xclang -target-feature -Xclang -packed-fp32-ops`` (packed FP32 VALU ops are
disabled on this hardware pool), ``-O3`` and ``--offload-arch=gfx950``.  The
library is written next to this file, where ``_lib.py`` loads it.
"""

from __future__ import annotations

import argparse
import concurrent.futures
import hashlib
import os
import shutil
import subprocess
import sys
from pathlib import Path

PACKAGE = Path(__file__).resolve().parent
CSRC = PACKAGE / "csrc"
INCLUDE = PACKAGE.parent / "include"
BUILD = PACKAGE / "_build"
LIBRARY = PACKAGE / "libtsn.so"
CLI = PACKAGE / "tsn_cg"
ARCH = os.environ.get("TSN_OFFLOAD_ARCH", "gfx950")

SOURCES = ["pack_io.cpp", "packer.cpp", "interp.hip", "fft.hip", "plan.hip", "batch.hip", "fft_compat.hip"]
COMMON = ["-O3", "-std=c++17", "-fPIC", f"--offload-arch={ARCH}", f"-I{INCLUDE}",
          "-Wall", "-Wno-unused-function"]
DEVICE = ["-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops"]


def hipcc() -> str:
    path = shutil.which("hipcc") or "/opt/rocm/bin/hipcc"
    if not os.path.exists(path):
        raise RuntimeError("hipcc not found; the HIP toolchain (ROCm) is required to build libtsn.so")
    return path


def _digest(paths, flags) -> str:
    h = hashlib.sha256(" ".join(flags).encode())
    for p in paths:
        h.update(Path(p).read_bytes())
    return h.hexdigest()[:16]


RESOURCE_REMARKS = ["-Rpass-analysis=kernel-resource-usage"]


def scratch_kernels(remarks: str) -> list[str]:
    """Kernels whose resource-usage remarks report private (scratch) memory:
    spilled registers or a local copy of a by-value argument, each a round
    trip to memory per access in a hot loop."""

    found, name = [], None
    for line in remarks.splitlines():
        if "remark: Function Name: " in line:
            name = line.split("Function Name: ", 1)[1].split(" [-R")[0].strip()
        elif "ScratchSize [bytes/lane]: " in line and name:
            if int(line.split("ScratchSize [bytes/lane]: ", 1)[1].split()[0]) > 0:
                found.append(name)
    return found


def _compile(source: str, force: bool) -> Path:
    src = CSRC / source
    headers = sorted(CSRC.glob("*.h")) + sorted(INCLUDE.glob("*.h"))
    hip_src = source.endswith(".hip")
    flags = COMMON + DEVICE + (["-x", "hip"] + RESOURCE_REMARKS if hip_src else [])
    obj = BUILD / f"{src.stem}-{_digest([src, *headers], flags)}.o"
    if obj.exists() and not force:
        return obj
    cmd = [hipcc(), *flags, "-c", str(src), "-o", str(obj) + ".tmp"]
    proc = subprocess.run(cmd, capture_output=True, text=True)
    noise = "'-packed-fp32-ops' is not a recognized feature"
    out = proc.stdout + proc.stderr
    # resource remarks (and the source excerpts printed under them) are parsed, not shown
    lines = [l for l in out.splitlines() if noise not in l]
    log = "\n".join(l for l in lines if "kernel-resource-usage" not in l and not l.startswith(" ")
                    and not l.startswith("In file included from"))
    if proc.returncode != 0:
        full = "\n".join(l for l in lines if "kernel-resource-usage" not in l)
        raise RuntimeError(f"hipcc failed for {source}:\n{full}")
    spills = scratch_kernels(out) if hip_src else []
    if spills:
        raise RuntimeError(f"{source}: kernels use scratch memory (spills or argument copies): {spills}")
    if log.strip():
        print(log, file=sys.stderr)
    os.replace(str(obj) + ".tmp", obj)
    return obj


def build(jobs: int | None = None, force: bool = True) -> Path:
    """Compile every source and link libtsn.so and the tsn_cg CLI."""

    BUILD.mkdir(exist_ok=True)
    jobs = jobs or min(8, os.cpu_count() or 1)
    with concurrent.futures.ThreadPoolExecutor(jobs) as pool:
        objects = list(pool.map(lambda s: _compile(s, force), SOURCES))
    tmp = LIBRARY.with_suffix(".so.tmp")
    cmd = [hipcc(), "-shared", f"--offload-arch={ARCH}", "-o", str(tmp), *map(str, objects), "-lpthread"]
    proc = subprocess.run(cmd, capture_output=True, text=True)
    if proc.returncode != 0:
        raise RuntimeError(f"link failed:\n{proc.stdout}{proc.stderr}")
    os.replace(tmp, LIBRARY)
    cli_src = CSRC / "tsn_cg_main.cpp"
    if cli_src.exists():
        cmd = [hipcc(), *COMMON, str(cli_src), "-o", str(CLI) + ".tmp", f"-L{PACKAGE}", "-ltsn",
               "-Wl,-rpath,$ORIGIN"]
        proc = subprocess.run(cmd, capture_output=True, text=True)
        if proc.returncode != 0:
            raise RuntimeError(f"CLI build failed:\n{proc.stdout}{proc.stderr}")
        os.replace(str(CLI) + ".tmp", CLI)
    return LIBRARY


def main() -> int:
    parser = argparse.ArgumentParser(description=__doc__)
    parser.add_argument("--jobs", type=int)
    parser.add_argument("--cached", action="store_true", help="reuse unchanged objects")
    args = parser.parse_args()
    print(build(args.jobs, force=not args.cached))
    return 0


if __name__ == "__main__":
    raise SystemExit(main())
