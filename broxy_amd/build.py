"""In-tree build of the GMorpho libraries.

libgmorpho_hip.so     HIP kernels + C-ABI (hipcc, gfx950)
libgmorpho_meshgen.so seeded workload generators (host C++)
libgmorpho_decimate.so proxy decimation of a closing mesh (host C++)
"""
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
LIBDIR = os.path.join(HERE, "_lib")

HIP_SOURCES = ["gm_context.hip", "gm_layout.hip", "gm_voxelize.hip", "gm_morph.hip", "gm_mc.hip",
               "gm_pyramid.hip", "gm_engine.hip", "gm_normals.hip",
               "gm_scale.hip", "gm_noalloc.hip", "gm_edt.hip"]
# gm_morph_ct.hip is compiled once per part (the compiled-radius kernels)
CT_SOURCE = "gm_morph_ct.hip"
CT_PARTS = 8
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = os.environ.get("GM_OFFLOAD_ARCH", "gfx950")

HIP_FLAGS = [
    "--offload-arch=" + ARCH, "-O3", "-std=c++17", "-fPIC", "-ffp-contract=off",
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


def build_decimator():
    """Host proxy decimation (csrc/decimator.cpp)."""
    os.makedirs(LIBDIR, exist_ok=True)
    out = os.path.join(LIBDIR, "libgmorpho_decimate.so")
    _run(["g++", "-O2", "-std=c++17", "-fPIC", "-shared", "-Wall", os.path.join(CSRC, "decimator.cpp"), "-o", out])
    return out


def build_hip(jobs=None):
    """Compiles every HIP source to an object (in parallel), then links."""
    os.makedirs(LIBDIR, exist_ok=True)
    objdir = os.path.join(LIBDIR, "obj")
    os.makedirs(objdir, exist_ok=True)
    units = [(os.path.join(CSRC, s), os.path.join(objdir, s.replace(".hip", ".o")), []) for s in HIP_SOURCES]
    for p in range(CT_PARTS):
        units.append((os.path.join(CSRC, CT_SOURCE), os.path.join(objdir, "gm_morph_ct%d.o" % p),
                      ["-DGM_CT_PART=%d" % p]))
    cmds = [[HIPCC] + HIP_FLAGS + extra + ["-c", src, "-o", obj] for src, obj, extra in units]
    jobs = jobs or min(len(cmds), max(1, os.cpu_count() or 1))
    running, failed = [], []
    pending = list(cmds)
    while pending or running:
        while pending and len(running) < jobs:
            cmd = pending.pop(0)
            print("+", " ".join(cmd), flush=True)
            running.append((cmd, subprocess.Popen(cmd)))
        cmd, proc = running.pop(0)
        if proc.wait() != 0:
            failed.append(cmd)
    if failed:
        raise subprocess.CalledProcessError(1, failed[0])
    out = os.path.join(LIBDIR, "libgmorpho_hip.so")
    tmp = out + ".tmp"
    _run([HIPCC, "--offload-arch=" + ARCH, "-shared", "-fPIC"] + [obj for _, obj, _ in units] + ["-o", tmp])
    os.replace(tmp, out)
    return out


def build_driver():
    """The C++ facade driver (tests/cpp/facade_closing.cpp), linked against the
    two libraries; run by the GPU tests."""
    bindir = os.path.join(LIBDIR, "bin")
    os.makedirs(bindir, exist_ok=True)
    out = os.path.join(bindir, "facade_closing")
    _run(["g++", "-O2", "-std=c++17", "-Wall", "-I", os.path.join(ROOT, "include"),
          os.path.join(ROOT, "tests", "cpp", "facade_closing.cpp"), "-o", out,
          "-L", LIBDIR, "-lgmorpho_hip", "-lgmorpho_meshgen", "-lgmorpho_decimate", "-L/opt/rocm/lib",
          "-Wl,-rpath,$ORIGIN/..", "-Wl,-rpath,/opt/rocm/lib"])
    return out


def build_decimate_driver():
    """The C++ Decimator facade driver (tests/cpp/decimate_facade.cpp), linked
    against the host decimation library only; run by the CPU tests."""
    bindir = os.path.join(LIBDIR, "bin")
    os.makedirs(bindir, exist_ok=True)
    out = os.path.join(bindir, "decimate_facade")
    _run(["g++", "-O2", "-std=c++17", "-Wall", "-I", os.path.join(ROOT, "include"),
          os.path.join(ROOT, "tests", "cpp", "decimate_facade.cpp"), "-o", out,
          "-L", LIBDIR, "-lgmorpho_decimate", "-Wl,-rpath,$ORIGIN/.."])
    return out


def build_oracle():
    _run(["make", "-C", os.path.join(ROOT, "oracle")])


def build_all():
    build_meshgen()
    build_decimator()
    build_decimate_driver()
    build_hip()
    build_driver()
    build_oracle()


if __name__ == "__main__":
    what = sys.argv[1:] or ["all"]
    for w in what:
        {"all": build_all, "hip": build_hip, "meshgen": build_meshgen, "decimator": build_decimator,
         "driver": build_driver, "decimate_driver": build_decimate_driver,
         "oracle": build_oracle}[w]()
