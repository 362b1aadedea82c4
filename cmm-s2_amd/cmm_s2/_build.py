"""Build libcmm_s2_amd.so from cmm-s2_amd/csrc/*.hip with hipcc for gfx950.

Objects go to cmm-s2_amd/_build/, the shared library next to this file.  Run
``python -m cmm_s2._build`` (with cmm-s2_amd on sys.path) or call build().
"""
import concurrent.futures as cf
import os
import subprocess
import sys

PKG = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(PKG)
CSRC = os.path.join(ROOT, "csrc")
OBJ = os.path.join(ROOT, "_build")
LIB = os.path.join(PKG, "libcmm_s2_amd.so")
ARCH = os.environ.get("CMM_OFFLOAD_ARCH", "gfx950")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")

CFLAGS = [
    "-O3", "-fPIC", "-std=c++17", f"--offload-arch={ARCH}",
    # packed FP32 instructions are disabled for every HIP translation unit
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-unused-result",
]


# Per-file flags.  mwss_mesh.hip evaluates the velocity spline in several
# kernel instantiations and rebuilds missing velocity blocks outside the
# assembly kernel; tri_mesh.hip locates and evaluates maps both point by
# point and along runs of points that reuse a located triangle.  With
# contraction decided per source expression (-ffp-contract=on: a * b + c
# fused as written, never across statements) every copy of a formula rounds
# the same way, so all of them give the same bits.
FILE_FLAGS = {"mwss_mesh.hip": ["-ffp-contract=on"], "tri_mesh.hip": ["-ffp-contract=on"]}


def sources():
    return sorted(os.path.join(CSRC, f) for f in os.listdir(CSRC) if f.endswith(".hip"))


def _compile(src):
    obj = os.path.join(OBJ, os.path.basename(src)[:-4] + ".o")
    cmd = [HIPCC, *CFLAGS, *FILE_FLAGS.get(os.path.basename(src), []), "-c", src, "-o", obj]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"hipcc failed for {src}:\n{r.stderr}")
    return obj


def build(verbose=False):
    os.makedirs(OBJ, exist_ok=True)
    srcs = sources()
    with cf.ThreadPoolExecutor(max_workers=min(8, len(srcs))) as ex:
        objs = list(ex.map(_compile, srcs))
    tmp = LIB + ".tmp"
    cmd = [HIPCC, "-shared", f"--offload-arch={ARCH}", "-o", tmp, *objs]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"link failed:\n{r.stderr}")
    os.replace(tmp, LIB)
    if verbose:
        print(f"built {LIB} from {len(srcs)} sources", file=sys.stderr)
    return LIB


if __name__ == "__main__":
    build(verbose=True)
