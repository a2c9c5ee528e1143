"""Build recipe for the gfx950 library (``gell_amd/_build/libgell_hip.so``).

Every ``csrc/*.hip`` file is compiled by hipcc for gfx950 only and linked into
one shared library that exports the C ABI of ``include/gell_abi.h``.

Flags that matter for results:
  -ffp-contract=off   no fused multiply-add, so a*b+c rounds twice exactly as
                      the reference's expressions are written
  (fp32 division and sqrt stay IEEE-correct: hipcc's default)
  -Xclang -target-feature -Xclang -packed-fp32-ops
                      no packed FP32 instructions
"""
from __future__ import annotations

import os
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor
from pathlib import Path

HERE = Path(__file__).resolve().parent
CSRC = HERE / "csrc"
OUT = HERE / "_build"
LIB = OUT / "libgell_hip.so"
ARCH = os.environ.get("GELL_OFFLOAD_ARCH", "gfx950")

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")

COMMON = [
    f"--offload-arch={ARCH}",
    "-O3",
    "-std=c++17",
    "-fPIC",
    "-ffp-contract=off",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall",
    "-Wno-unused-command-line-argument",
]


def sources() -> list[Path]:
    return sorted(CSRC.glob("*.hip"))


def _compile(src: Path) -> Path:
    obj = OUT / (src.stem + ".o")
    cmd = [HIPCC, *COMMON, "-c", str(src), "-o", str(obj)]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"hipcc failed on {src.name}:\n{r.stdout}\n{r.stderr}")
    return obj


def build(verbose: bool = False) -> Path:
    """Compile every HIP source and link the shared library; returns its path."""
    OUT.mkdir(exist_ok=True)
    srcs = sources()
    jobs = min(len(srcs), int(os.environ.get("MAX_JOBS", os.cpu_count() or 4)), 16)
    with ThreadPoolExecutor(max_workers=max(1, jobs)) as ex:
        objs = list(ex.map(_compile, srcs))
    tmp = LIB.with_suffix(".so.tmp")
    cmd = [HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", str(tmp), *map(str, objs)]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"link failed:\n{r.stdout}\n{r.stderr}")
    os.replace(tmp, LIB)
    # C++ driver with the reference's time-loop structure, linked to the library
    sim = OUT / "gell_sim"
    cmd = [HIPCC, f"--offload-arch={ARCH}", "-O2", "-std=c++17", "-o", str(sim), str(CSRC / "gell_sim.cpp"),
           f"-L{OUT}", "-lgell_hip", "-Wl,-rpath,$ORIGIN"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"gell_sim build failed:\n{r.stdout}\n{r.stderr}")
    if verbose:
        print(f"built {LIB} and {sim}")
    return LIB


if __name__ == "__main__":
    build(verbose=True)
    sys.exit(0)
