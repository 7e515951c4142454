"""Build the native library ``libdiffsims_amd.so`` in-tree with hipcc (gfx950).

The library is built from the HIP sources in ``diffsims_amd/csrc`` with the C
ABI declared in ``include/diffsims_amd.h``.  Usage::

    python -m diffsims_amd.build
"""

import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
INCLUDE = os.path.join(ROOT, "include")
LIB_NAME = "libdiffsims_amd.so"
LIB_PATH = os.path.join(HERE, LIB_NAME)

ARCH = "gfx950"
FLAGS = [
    "-O3",
    f"--offload-arch={ARCH}",
    "-std=c++17",
    "-fPIC",
    "-shared",
    # keep FP32 arithmetic scalar: no packed v_pk_*_f32 instructions
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-unused-command-line-argument",
]


def sources():
    return sorted(os.path.join(CSRC, f) for f in os.listdir(CSRC) if f.endswith(".hip"))


def hipcc():
    for cand in (shutil.which("hipcc"), "/opt/rocm/bin/hipcc"):
        if cand and os.path.exists(cand):
            return cand
    raise RuntimeError("hipcc not found (ROCm toolchain required to build diffsims_amd)")


def build(verbose: bool = False) -> str:
    """Compile every HIP source into LIB_PATH (always recompiles)."""
    tmp = LIB_PATH + ".tmp"
    cmd = [hipcc(), *FLAGS, f"-I{INCLUDE}", "-o", tmp, *sources()]
    if verbose:
        print(" ".join(cmd), flush=True)
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        sys.stderr.write(res.stdout + res.stderr)
        raise RuntimeError(f"hipcc failed with exit code {res.returncode}")
    os.replace(tmp, LIB_PATH)
    return LIB_PATH


if __name__ == "__main__":
    print(build(verbose=True))
