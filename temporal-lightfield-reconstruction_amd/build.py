"""Builds the native libraries in-tree.

lib/libtlr.so        HIP kernels + C ABI (hipcc, gfx950)
lib/libtlr_scene.so  synthetic sample-buffer generator (g++, host only)

FP rules: no contraction (-ffp-contract=off), IEEE division and square root
(HIP's default correctly-rounded f32 div/sqrt), no fast-math, so device
arithmetic matches an IEEE restatement of the reference operation for
operation.  Packed FP32 instructions are disabled.
"""
import os
import subprocess
import sys

PKG = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(PKG)
CSRC = os.path.join(PKG, "csrc")
LIB = os.path.join(PKG, "lib")

HIP_SOURCES = ["prims.hip", "build.hip", "query.hip", "refgpu.hip", "gate.hip", "noalloc.hip", "shadow.hip", "tlr.cpp",
               "pattern.cpp", "sample_io.cpp", "shadow_setup.cpp"]
HIP_FLAGS = [
    "--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC", "-shared",
    "-ffp-contract=off", "-fno-fast-math",
    # the SLP vectorizer turns the node tests' four-compare AND into a
    # vector compare plus a 16-bit bit reduction (about 12 more VALU
    # instructions per node visited than three scalar mask ANDs)
    "-fno-slp-vectorize",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-unused-result",
]


def _run(cmd):
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stdout + r.stderr)
        raise RuntimeError("build failed: " + " ".join(cmd[:3]))
    return r


def build_scene():
    os.makedirs(LIB, exist_ok=True)
    out = os.path.join(LIB, "libtlr_scene.so")
    _run(["g++", "-O2", "-std=c++17", "-fPIC", "-shared", "-o", out,
          os.path.join(CSRC, "scene", "scene.cpp"), "-lpthread"])
    return out


def build_tlr(name="libtlr.so", defines=()):
    os.makedirs(LIB, exist_ok=True)
    out = os.path.join(LIB, name)
    tmp = out + ".tmp"
    srcs = [os.path.join(CSRC, s) for s in HIP_SOURCES]
    _run(["hipcc"] + HIP_FLAGS + ["-D" + d for d in defines] + ["-I", os.path.join(ROOT, "include"), "-o", tmp] +
         srcs)
    os.replace(tmp, out)
    return out


def build_tlr_small_lds():
    """Test build: the general kernel's LDS leaf staging and frontier stack at
    their smallest, so that the paths past them run on every query, and the
    1R lane passes' per-lane leaf lists small (4, then 8), so that ordinary
    scenes take both lane passes and the general kernel
    (tests/test_gpu_small_lds.py)."""
    return build_tlr("libtlr_small_lds.so", ("TLR_GATHER_STAGE=4", "TLR_GATHER_LDS_STACK=64",
                                             "TLR_FAST_LEAF_CAP=4", "TLR_WIDE_LEAF_CAP=8"))


def build_tlr_tiny_cap():
    """Test build: the general kernel holds at most 8 leaves per output
    sample, so ordinary scenes reach the TLR_E_CAPACITY error path
    (tests/test_gpu_edges.py)."""
    return build_tlr("libtlr_tiny_cap.so", ("TLR_LEAF_CAP=8",))


def build_facade_example(name="facade_example"):
    """C++ caller of the facade (tests/cpp/<name>.cpp), host compiler only."""
    out = os.path.join(LIB, name)
    _run(["g++", "-O2", "-std=c++17", "-I", os.path.join(ROOT, "include"), "-I", CSRC,
          "-o", out, os.path.join(ROOT, "tests", "cpp", name + ".cpp"),
          "-L", LIB, "-ltlr", "-Wl,-rpath,$ORIGIN"])
    return out


def build_all():
    return [build_scene(), build_tlr(), build_tlr_small_lds(), build_tlr_tiny_cap(), build_facade_example(),
            build_facade_example("facade_shadow_example")]


if __name__ == "__main__":
    for p in build_all():
        print(p)
