"""Builds libaugur_hip.so (gfx950) in-tree with hipcc.

Each ``csrc/*.hip`` file is compiled to an object under ``_build/`` and the
objects are linked into ``augurv2_amd/libaugur_hip.so``.
"""
import os
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
BUILD = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "libaugur_hip.so")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")

ARCH = "gfx950"
FLAGS = [
    f"--offload-arch={ARCH}", "-O3", "-fPIC", "-std=c++17",
    "-Wall", "-Wno-unused-function", "-Wno-unused-variable",
    # packed FP32 VALU ops stay off: the compiler may form them from scalar
    # float code, and some of their forms have produced wrong results on
    # the target pool
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-munsafe-fp-atomics",
    "-I", os.path.join(HERE, "..", "include"),
]


def sources():
    return sorted(os.path.join(CSRC, f) for f in os.listdir(CSRC) if f.endswith(".hip"))


# Per-file flags: matrix-core results in ordinary VGPRs rather than
# accumulation registers.  augur_gmm.hip: its z step otherwise copies every
# product back with v_accvgpr_read (512 instructions per 64 points);
# augur_bhlr.hip: the matrix-core data pass drops from 161 + 36 registers
# (two waves per SIMD) to 164 (three) and loses its 243 copies.
_VGPR_MFMA = ["-mllvm", "-amdgpu-mfma-vgpr-form=1"]
FILE_FLAGS = {"augur_gmm.hip": _VGPR_MFMA, "augur_bhlr.hip": _VGPR_MFMA}


def _compile(src):
    obj = os.path.join(BUILD, os.path.basename(src)[:-4] + ".o")
    cmd = [HIPCC] + FLAGS + FILE_FLAGS.get(os.path.basename(src), []) + ["-c", src, "-o", obj]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"hipcc failed for {src}:\n{r.stderr}")
    return obj


def build(verbose=False):
    os.makedirs(BUILD, exist_ok=True)
    srcs = sources()
    with ThreadPoolExecutor(max_workers=min(8, len(srcs))) as ex:
        objs = list(ex.map(_compile, srcs))
    tmp = LIB + ".tmp"
    # hipRTC compiles the passes of generic models at run time
    cmd = [HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", tmp] + objs + [
        "-L/opt/rocm/lib", "-lhiprtc", "-Wl,-rpath,/opt/rocm/lib"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"link failed:\n{r.stderr}")
    os.replace(tmp, LIB)
    if verbose:
        print("built", LIB)
    return LIB


if __name__ == "__main__":
    build(verbose=True)
    sys.exit(0)
