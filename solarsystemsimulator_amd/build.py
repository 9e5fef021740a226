"""Build libssim.so (hand-written HIP for gfx950) in-tree with hipcc.

Run as ``python -m solarsystemsimulator_amd.build``.  Every ``csrc/*.hip``
file is compiled to an object in ``_build/`` (in parallel) and linked into
``solarsystemsimulator_amd/libssim.so``, which the ctypes loader in
``_lib.py`` opens.  hipcc cross-compiles for gfx950 without a GPU present.
"""
from __future__ import annotations

import concurrent.futures as cf
import os
import subprocess
import sys
from pathlib import Path

PKG = Path(__file__).resolve().parent
CSRC = PKG / "csrc"
BUILD = PKG.parent / "_build"
LIB = PKG / "libssim.so"
ARCH = "gfx950"
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")

# No FMA contraction: the integrator kernels must round exactly like the
# reference's NumPy updates; every fma in the force loops is written out.
# Packed FP32 instructions are switched off for the device code.
CFLAGS = [
    f"--offload-arch={ARCH}", "-O3", "-std=c++17", "-fPIC",
    "-ffp-contract=off",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-function", "-Wno-unused-command-line-argument",
]


# Per-file code generation options: MFMA results in VGPRs rather than AGPRs,
# so the near field's filter (pp.hip) reads its r^2 tiles and the network's
# activations (mlp.hip) read their accumulators without v_accvgpr_read copies.
_VGPR_FORM = ["-mllvm", "-amdgpu-mfma-vgpr-form=1"]
FILE_FLAGS = {"pp.hip": _VGPR_FORM, "mlp.hip": _VGPR_FORM}


def _compile(src: Path) -> Path:
    obj = BUILD / (src.stem + ".o")
    cmd = [HIPCC, "-x", "hip", "-c", str(src), "-o", str(obj), *CFLAGS,
           *FILE_FLAGS.get(src.name, [])]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"hipcc failed for {src.name}:\n{r.stdout}\n{r.stderr}")
    return obj


def build(verbose: bool = True) -> Path:
    """Compile every HIP source and link libssim.so; returns its path."""
    BUILD.mkdir(exist_ok=True)
    srcs = sorted(CSRC.glob("*.hip"))
    if not srcs:
        raise RuntimeError(f"no HIP sources under {CSRC}")
    workers = min(len(srcs), max(1, (os.cpu_count() or 2) - 1), 8)
    with cf.ThreadPoolExecutor(workers) as ex:
        objs = list(ex.map(_compile, srcs))
    tmp = LIB.with_suffix(".so.tmp")
    cmd = [HIPCC, "-shared", "-fPIC", f"--offload-arch={ARCH}", "-o", str(tmp),
           *map(str, objs)]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"link failed:\n{r.stdout}\n{r.stderr}")
    os.replace(tmp, LIB)
    if verbose:
        print(f"built {LIB} from {len(srcs)} sources", file=sys.stderr)
    return LIB


if __name__ == "__main__":
    build()
