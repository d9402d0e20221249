"""Build libtisphi_hip.so for gfx950 with hipcc (in-tree, no JIT cache)."""
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
LIB = os.path.join(HERE, "libtisphi_hip.so")
SOURCES = [os.path.join(CSRC, "tsph_engine.hip")]
DEPS = SOURCES + [os.path.join(CSRC, f) for f in ("tsph_device.h", "tsph_kernels.h")] + [
    os.path.join(HERE, "..", "include", "tisphi.h")]

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
FLAGS = [
    "--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC", "-shared",
    # Packed FP32 (v_pk_*_f32) off: plain scalar VALU FP32 only.
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    # No implicit mul+add contraction: every product and sum rounds as
    # written (fmaf only where the source says so), so one formula gives the
    # same bits in every kernel that inlines it and in the oracle.
    "-ffp-contract=off",
    "-Wno-unused-result",
]


def build(verbose=False, extra=()):
    """Compile the engine library; always recompiles."""
    cmd = [HIPCC, *FLAGS, *extra, "-o", LIB + ".tmp", *SOURCES]
    if verbose:
        print(" ".join(cmd), flush=True)
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stdout + r.stderr)
        raise RuntimeError("hipcc failed")
    os.replace(LIB + ".tmp", LIB)
    return LIB


if __name__ == "__main__":
    print(build(verbose=True))
