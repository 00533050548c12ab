"""Build libfv3d_hip.so (HIP for gfx950) in-tree.

Compiles every source under fv3d_amd/csrc with hipcc for gfx950 and links
fv3d_amd/libfv3d_hip.so.  The parity build flags are part of the product:
-ffp-contract=off keeps the FP64 arithmetic operation-for-operation equal to
the reference's x86-64 build (no fused multiply-adds).
"""
import os
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
BUILD = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "libfv3d_hip.so")

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
# host-only translation units: the ROCm clang in plain C++ mode
HOSTCXX = os.environ.get("FVH_HOSTCXX", "/opt/rocm/lib/llvm/bin/clang++")
ARCH = "gfx950"
COMMON = [
    "-O3", "-std=c++17", "-fPIC", "-ffp-contract=off",
    "-I" + os.path.join(ROOT, "include"), "-I" + CSRC,
    "-Wall", "-Wno-unused-result",
]
HIP_ONLY = [
    "--offload-arch=" + ARCH,
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
]


def _sources():
    return sorted(
        f for f in os.listdir(CSRC) if f.endswith((".hip", ".cpp"))
    )


def _compile(src):
    obj = os.path.join(BUILD, src + ".o")
    if src.endswith(".hip"):
        cmd = [HIPCC] + COMMON + HIP_ONLY + ["-x", "hip"]
    else:
        cmd = [HOSTCXX] + COMMON + ["-D__HIP_PLATFORM_AMD__", "-I/opt/rocm/include"]
    cmd += ["-c", os.path.join(CSRC, src), "-o", obj]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("compile failed: %s\n%s\n%s" % (" ".join(cmd), r.stdout, r.stderr))
    return obj


def build(verbose=False):
    os.makedirs(BUILD, exist_ok=True)
    srcs = _sources()
    with ThreadPoolExecutor(max_workers=min(8, len(srcs))) as ex:
        objs = list(ex.map(_compile, srcs))
    tmp = LIB + ".tmp"
    cmd = [HIPCC, "--offload-arch=" + ARCH, "-shared", "-fPIC", "-o", tmp] + objs + ["-ldl"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("link failed: %s\n%s\n%s" % (" ".join(cmd), r.stdout, r.stderr))
    os.replace(tmp, LIB)
    if verbose:
        print("built", LIB)
    return LIB


if __name__ == "__main__":
    build(verbose=True)
    sys.exit(0)
