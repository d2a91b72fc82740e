"""Build libmfgpu.so (HIP, gfx950) in-tree.

Each source is compiled with hipcc into an object under _build/, then linked
into matrix-free-dealii-precice_amd/libmfgpu.so.  Usage: python build.py [-j N]
"""
import concurrent.futures as cf
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
BUILD = os.path.join(HERE, "_build")
OUT = os.path.join(HERE, "libmfgpu.so")

SOURCES = ["setup.cpp", "operator.hip", "vectors.hip", "multigrid.hip", "newton.hip", "batched.hip", "serving.cpp",
           "heat_problem.hip", "interface.hip", "cg_noalloc.hip"]

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
# -packed-fp32-ops: keep the compiler from forming v_pk_* FP32 instructions.
# -ffp-contract=on: a multiply and an add fuse only within one source
# expression, so that the rounding of a kernel's arithmetic follows from its
# source alone and not from the code around it (the batched kernels and the
# single-operator kernels, the colour and the GATHER schedule share their
# arithmetic and stay bitwise equal).
FLAGS = ["-O3", "-std=c++17", "-fPIC", "--offload-arch=gfx950", "-ffp-contract=on", "-munsafe-fp-atomics",
         "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
         "-Wno-unused-result", "-I", os.path.join(HERE, "..", "include")]


def _compile(src):
    path = os.path.join(CSRC, src)
    obj = os.path.join(BUILD, os.path.splitext(src)[0] + ".o")
    lang = ["-x", "hip"] if src.endswith(".hip") else []
    cmd = [HIPCC] + FLAGS + lang + ["-c", path, "-o", obj]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"compile failed: {src}\n{' '.join(cmd)}\n{r.stdout}\n{r.stderr}")
    return obj


def build(jobs=8, verbose=True):
    os.makedirs(BUILD, exist_ok=True)
    sources = [s for s in SOURCES if os.path.exists(os.path.join(CSRC, s))]
    with cf.ThreadPoolExecutor(max_workers=jobs) as ex:
        objs = list(ex.map(_compile, sources))
    cmd = [HIPCC, "-shared", "-fPIC", "--offload-arch=gfx950", "-o", OUT + ".tmp"] + objs
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"link failed\n{r.stdout}\n{r.stderr}")
    os.replace(OUT + ".tmp", OUT)
    # C++ drop-in test program against the host mirror (include/mfgpu.hpp)
    src = os.path.join(HERE, "..", "tests", "cpp", "drop_in.cpp")
    if os.path.exists(src):
        exe = os.path.join(BUILD, "drop_in_test")
        cmd = [HIPCC, "-O2", "-std=c++17", "-Wno-unused-result", src, "-L", HERE, "-lmfgpu",
               "-Wl,-rpath,$ORIGIN/..", "-o", exe]
        r = subprocess.run(cmd, capture_output=True, text=True)
        if r.returncode != 0:
            raise RuntimeError(f"drop-in test build failed\n{r.stdout}\n{r.stderr}")
    if verbose:
        print(f"built {OUT}")
    return OUT


if __name__ == "__main__":
    j = int(sys.argv[sys.argv.index("-j") + 1]) if "-j" in sys.argv else 8
    build(j)
