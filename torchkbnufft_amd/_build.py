"""Compile ``libtkbn_hip.so`` for gfx950 with hipcc (explicit, in-tree).

``python -m torchkbnufft_amd._build`` or ``__graft_entry__.build()``.

Flags:
* ``--offload-arch=gfx950``: MI355X only.
* ``-O3`` without fast-math: the index arithmetic needs IEEE float32 division,
  ``rint`` (half-to-even) and ``floor`` exactly as the reference evaluates them.
* ``-munsafe-fp-atomics``: hardware float atomics for the atomic adjoint
  (device memory from the torch caching allocator is coarse-grained).
* ``-Xclang -target-feature -Xclang -packed-fp32-ops``: no packed FP32 VALU ops.
"""
from __future__ import annotations

import concurrent.futures as cf
import os
import shutil
import subprocess
import sys

PKG = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(PKG)
CSRC = os.path.join(PKG, "csrc")
BUILD = os.path.join(PKG, "_build")
LIB = os.path.join(PKG, "libtkbn_hip.so")

SOURCES = ["tkbn_interp.hip", "tkbn_spread.hip", "tkbn_fftscale.hip", "tkbn_fft.hip", "tkbn_spmat.hip"]

HIPCC_FLAGS = [
    "--offload-arch=gfx950",
    "-O3",
    "-std=c++17",
    "-fPIC",
    "-fvisibility=hidden",
    "-munsafe-fp-atomics",
    "-Xclang",
    "-target-feature",
    "-Xclang",
    "-packed-fp32-ops",
    "-Wno-unused-command-line-argument",
    "-I" + os.path.join(ROOT, "include"),
]


def _hipcc() -> str:
    for cand in (shutil.which("hipcc"), "/opt/rocm/bin/hipcc"):
        if cand and os.path.exists(cand):
            return cand
    raise RuntimeError("hipcc not found")


def _torch_lib_dir() -> str:
    import torch  # noqa: F401  (only to locate its bundled HIP runtime)

    return os.path.join(os.path.dirname(torch.__file__), "lib")


def build(verbose: bool = False, jobs: int = 4) -> str:
    os.makedirs(BUILD, exist_ok=True)
    hipcc = _hipcc()

    def compile_one(src: str) -> str:
        obj = os.path.join(BUILD, os.path.splitext(src)[0] + ".o")
        cmd = [hipcc, *HIPCC_FLAGS, "-c", os.path.join(CSRC, src), "-o", obj]
        if verbose:
            print(" ".join(cmd), flush=True)
        r = subprocess.run(cmd, capture_output=True, text=True)
        if r.returncode != 0:
            raise RuntimeError(f"hipcc failed on {src}:\n{r.stdout}\n{r.stderr}")
        return obj

    with cf.ThreadPoolExecutor(max_workers=jobs) as ex:
        objs = list(ex.map(compile_one, SOURCES))
    tmp = LIB + ".tmp"
    # link against torch's bundled HIP runtime (same SONAME libamdhip64.so.7 that
    # torch has already loaded when the library is opened after `import torch`)
    cmd = [hipcc, "--offload-arch=gfx950", "-shared", "-fPIC", *objs, "-o", tmp,
           "-L" + _torch_lib_dir(), "-lamdhip64"]
    if verbose:
        print(" ".join(cmd), flush=True)
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"link failed:\n{r.stdout}\n{r.stderr}")
    os.replace(tmp, LIB)
    return LIB


if __name__ == "__main__":
    print(build(verbose="-v" in sys.argv))
