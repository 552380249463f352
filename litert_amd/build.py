"""Build recipe for liblrt_hip.so (HIP kernels for gfx950 + C++ host code).

Run ``python -m litert_amd.build`` (or ``__graft_entry__.build()``). Host C++
is compiled with g++ (OpenMP); HIP sources with hipcc for gfx950 only; the
shared library is linked in-tree at litert_amd/liblrt_hip.so.
"""
import os
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
BUILD = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "liblrt_hip.so")
ROCM = os.environ.get("ROCM_PATH", "/opt/rocm")

HOST_SOURCES = ["host/scene_gen.cpp", "host/bvh_build.cpp", "host/hydra_scene.cpp"]
HIP_SOURCES = ["lrt_capi.hip", "lrt_dr_capi.hip", "lrt_mlp_capi.hip", "lrt_build_capi.hip", "lrt_noalloc.hip", "kernels/render.hip",
               "kernels/dr.hip", "kernels/mlp.hip", "kernels/build.hip"]

# -ffp-contract=off: kernels evaluate every expression in the reference's
# operation order, so the FP32 results can be compared bit for bit with a CPU
# FP32 evaluation. Packed FP32 instructions are disabled: some of their forms
# have produced wrong results on this hardware without any error.
HIP_FLAGS = [
    "--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC", "-ffp-contract=off",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-munsafe-fp-atomics", "-Wno-unused-result",
]
HOST_FLAGS = ["-O3", "-std=c++17", "-fPIC", "-fopenmp", "-ffp-contract=off", "-Wall", "-Wno-unused-function"]


def _run(cmd):
    r = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    if r.returncode != 0:
        raise RuntimeError("build step failed:\n" + " ".join(cmd) + "\n" + r.stdout)
    return r.stdout


def build(verbose=False, jobs=None, extra_flags=(), out=None, build_dir=None):
    """Compiles the HIP library; extra_flags / out / build_dir allow variant builds."""
    bdir = build_dir or BUILD
    os.makedirs(bdir, exist_ok=True)
    hipcc = os.path.join(ROCM, "bin", "hipcc")
    cmds = []
    objs = []
    for src in HOST_SOURCES:
        obj = os.path.join(bdir, os.path.basename(src) + ".o")
        cmds.append(["g++", *HOST_FLAGS, "-c", os.path.join(CSRC, src), "-o", obj])
        objs.append(obj)
    for src in HIP_SOURCES:
        path = os.path.join(CSRC, src)
        obj = os.path.join(bdir, os.path.basename(src) + ".o")
        cmds.append([hipcc, *HIP_FLAGS, *extra_flags, "-c", path, "-o", obj])
        objs.append(obj)
    jobs = jobs or min(8, os.cpu_count() or 4)
    with ThreadPoolExecutor(max_workers=jobs) as ex:
        outs = list(ex.map(_run, cmds))
    if verbose:
        for c, o in zip(cmds, outs):
            print(" ".join(c))
            if o.strip():
                print(o)
    lib = out or LIB
    tmp = lib + ".tmp"
    _run(["g++", "-shared", "-fopenmp", "-o", tmp, *objs, "-L" + os.path.join(ROCM, "lib"), "-lamdhip64",
          "-Wl,-rpath," + os.path.join(ROCM, "lib"), "-Wl,--no-undefined"])
    os.replace(tmp, lib)
    return lib


def build_cpp_tests(oracle_lib):
    """C++ tests that use the facade header (include/litert_amd/MultiRenderer.hpp)."""
    out_dir = os.path.join(ROOT, "tests", "cpp", "_build")
    os.makedirs(out_dir, exist_ok=True)
    exes = []
    for name in ("test_facade_render", "test_facade_dr"):
        exe = os.path.join(out_dir, name)
        _run(["g++", "-O2", "-std=c++17", "-Wall", os.path.join(ROOT, "tests", "cpp", name + ".cpp"), "-o", exe,
              LIB, oracle_lib, "-Wl,-rpath," + HERE, "-Wl,-rpath," + os.path.dirname(oracle_lib), "-fopenmp"])
        exes.append(exe)
    return exes


if __name__ == "__main__":
    print(build(verbose="-v" in sys.argv))
