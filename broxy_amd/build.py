"""In-tree build of the GMorpho libraries.

libgmorpho_hip.so     HIP kernels + C-ABI (hipcc, gfx950)
libgmorpho_meshgen.so seeded workload generators (host C++)
"""
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
LIBDIR = os.path.join(HERE, "_lib")

HIP_SOURCES = ["gm_context.hip", "gm_layout.hip", "gm_voxelize.hip", "gm_morph.hip", "gm_mc.hip",
               "gm_pyramid.hip"]
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = os.environ.get("GM_OFFLOAD_ARCH", "gfx950")

HIP_FLAGS = [
    "--offload-arch=" + ARCH, "-O3", "-std=c++17", "-fPIC", "-shared", "-ffp-contract=off",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-function", "-Wno-unused-variable", "-Wno-unused-value", "-Wno-unused-result",
]


def _run(cmd, cwd=None):
    print("+", " ".join(cmd), flush=True)
    subprocess.run(cmd, check=True, cwd=cwd)


def build_meshgen():
    os.makedirs(LIBDIR, exist_ok=True)
    out = os.path.join(LIBDIR, "libgmorpho_meshgen.so")
    _run(["g++", "-O2", "-std=c++17", "-fPIC", "-shared", "-ffp-contract=off",
          os.path.join(CSRC, "meshgen.cpp"), "-o", out])
    return out


def build_hip():
    os.makedirs(LIBDIR, exist_ok=True)
    out = os.path.join(LIBDIR, "libgmorpho_hip.so")
    tmp = out + ".tmp"
    _run([HIPCC] + HIP_FLAGS + [os.path.join(CSRC, s) for s in HIP_SOURCES] + ["-o", tmp])
    os.replace(tmp, out)
    return out


def build_oracle():
    _run(["make", "-C", os.path.join(ROOT, "oracle")])


def build_all():
    build_meshgen()
    build_hip()
    build_oracle()


if __name__ == "__main__":
    what = sys.argv[1:] or ["all"]
    for w in what:
        {"all": build_all, "hip": build_hip, "meshgen": build_meshgen, "oracle": build_oracle}[w]()
