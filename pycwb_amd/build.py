"""In-tree build of libpycwb_amd.so with hipcc for gfx950.

Flags that the bit-exact parity with the float reference depends on:
  -ffp-contract=off                 no FMA contraction of a*b+c
  -fno-gpu-flush-denormals-to-zero  IEEE denormals (the default, made explicit)
  (default) correctly rounded f32 divide and sqrt
  -target-feature -packed-fp32-ops  no packed FP32 instructions
"""

from __future__ import annotations

import glob
import os
import subprocess

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
LIB = os.path.join(HERE, "libpycwb_amd.so")

HIPCC_FLAGS = [
    "--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC", "-shared",
    "-ffp-contract=off", "-fno-gpu-flush-denormals-to-zero",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-unused-result", "-Wno-unused-value",
]


def sources() -> list[str]:
    return sorted(glob.glob(os.path.join(HERE, "csrc", "*.hip")))


def _compile(hipcc: str, src: str, obj: str, verbose: bool) -> None:
    flags = [f for f in HIPCC_FLAGS if f != "-shared"]
    cmd = [hipcc, *flags, "-c", "-I", os.path.join(ROOT, "include"), "-I",
           os.path.join(HERE, "csrc"), "-o", obj, src]
    if verbose:
        print(" ".join(cmd), flush=True)
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"hipcc failed on {os.path.basename(src)} ({r.returncode}):\n"
                           f"{r.stderr[-4000:]}")


def build(verbose: bool = False, jobs: int | None = None) -> str:
    """Compiles every csrc/*.hip to an object (in parallel) and links the
    shared library."""
    from concurrent.futures import ThreadPoolExecutor

    hipcc = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
    objdir = os.path.join(ROOT, "_build", "obj")
    os.makedirs(objdir, exist_ok=True)
    srcs = sources()
    objs = [os.path.join(objdir, os.path.basename(s)[:-4] + ".o") for s in srcs]
    jobs = jobs or max(1, min(len(srcs), (os.cpu_count() or 4)))
    with ThreadPoolExecutor(max_workers=jobs) as ex:
        for f in [ex.submit(_compile, hipcc, s, o, verbose) for s, o in zip(srcs, objs)]:
            f.result()
    tmp = LIB + ".tmp"
    cmd = [hipcc, "--offload-arch=gfx950", "-shared", "-fPIC", "-o", tmp, *objs]
    if verbose:
        print(" ".join(cmd), flush=True)
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"hipcc link failed ({r.returncode}):\n{r.stderr[-4000:]}")
    os.replace(tmp, LIB)
    return LIB


if __name__ == "__main__":
    print(build(verbose=True))
