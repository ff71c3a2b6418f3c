"""Build libabc.so (gfx950) in-tree with hipcc.

    python -m abc_inference_transcription_amd.build

Every floating-point operation is compiled as written (-ffp-contract=off): the
GPU results must equal a scalar CPU replay bit for bit.  Packed FP32 VALU ops
are disabled for the whole library.
"""
from __future__ import annotations

import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
LIB = os.path.join(HERE, "libabc.so")
SOURCES = ["abc_kernels.hip", "abc_capi.hip", "abc_gemm.hip", "abc_accept.hip", "abc_mlp.hip", "abc_moments.hip", "abc_select.hip", "abc_data.hip"]
HEADERS = ["abc_device.h", "abc_kernels.h", "abc_logtab.h"]

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
FLAGS = [
    "--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC", "-shared",
    "-ffp-contract=off", "-fno-fast-math",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-unused-result",
]


def sources():
    return [os.path.join(CSRC, s) for s in SOURCES if os.path.exists(os.path.join(CSRC, s))]


def build(verbose: bool = False) -> str:
    """Compile every HIP source into libabc.so (always recompiles)."""
    tmp = LIB + ".tmp"
    cmd = [HIPCC, *FLAGS, "-I", os.path.join(os.path.dirname(HERE), "include"), "-o", tmp,
           *sources()]
    if verbose:
        print(" ".join(cmd), flush=True)
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stdout + r.stderr)
        raise RuntimeError("hipcc failed building libabc.so")
    os.replace(tmp, LIB)
    return LIB


if __name__ == "__main__":
    print(build(verbose=True))
