"""Build the HIP library sphinxsys_amd/lib/libsphx.so for gfx950 (in-tree) and the
C++ programs that drive it: the dam-break case (sphinxsys_amd/lib/dambreak) and
the neighbour-search microbenchmark (sphinxsys_amd/lib/microbench)."""
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
LIB_DIR = os.path.join(HERE, "lib")
LIB = os.path.join(LIB_DIR, "libsphx.so")
OBJ_DIR = os.path.join(HERE, "_build")

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
CXX = os.environ.get("CXX", "g++")
CASE_EXE = os.path.join(LIB_DIR, "dambreak")
MICROBENCH_EXE = os.path.join(LIB_DIR, "microbench")
ARCH = "gfx950"
COMMON = ["-O3", "-std=c++17", "-fPIC", f"--offload-arch={ARCH}",
          # packed FP32 instructions stay off: the compiler must not form v_pk_* FP32 ops
          "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops"]
# index/criterion work is compared bit-exactly with the CPU path: no contraction there
SOURCES = {
    "config_kernels.hip": ["-ffp-contract=off"],
    "fluid_kernels.hip": [],
    "sphx_api.hip": [],
    "batch.hip": [],
}


def _run(cmd, verbose):
    if verbose:
        print(" ".join(cmd), flush=True)
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stdout + r.stderr)
        raise RuntimeError(f"build step failed: {' '.join(cmd[:3])} ...")


def build(verbose=False):
    os.makedirs(LIB_DIR, exist_ok=True)
    os.makedirs(OBJ_DIR, exist_ok=True)
    # the translation units compile side by side
    from concurrent.futures import ThreadPoolExecutor
    jobs = []
    for src, extra in SOURCES.items():
        obj = os.path.join(OBJ_DIR, src.replace(".hip", ".o"))
        jobs.append(([HIPCC, *COMMON, *extra, "-c", os.path.join(CSRC, src), "-o", obj], obj))
    with ThreadPoolExecutor(max_workers=len(jobs)) as pool:
        for f in [pool.submit(_run, cmd, verbose) for cmd, _ in jobs]:
            f.result()
    objs = [obj for _, obj in jobs]
    tmp = LIB + ".tmp"
    _run([HIPCC, *COMMON, "-shared", "-o", tmp, *objs], verbose)
    os.replace(tmp, LIB)
    # the C++ programs on the CK mirror (host code only: they load LIB at run time)
    for exe, src in ((CASE_EXE, "dambreak.cpp"), (MICROBENCH_EXE, "microbench.cpp")):
        _run([CXX, "-O2", "-std=c++17", "-Wall", "-o", exe + ".tmp", os.path.join(HERE, "host", src), "-ldl"],
             verbose)
        os.replace(exe + ".tmp", exe)
    return LIB


if __name__ == "__main__":
    print(build(verbose=True))
