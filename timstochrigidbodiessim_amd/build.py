"""Build libfcm_hip.so in-tree with hipcc for gfx950.

    python -m timstochrigidbodiessim_amd.build [--jobs N]

Each translation unit is compiled in parallel into _build/, then linked into
timstochrigidbodiessim_amd/libfcm_hip.so.  All HIP code is compiled with packed
FP32 instructions disabled.
"""
import argparse
import concurrent.futures as cf
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
BUILD = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "libfcm_hip.so")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")

SOURCES = ["fcm_fft.hip", "fcm_grid.hip", "fcm_solver.hip", "fcm_gmres.hip", "fcm_capi.hip", "fcm_batch.hip", "fcm_steric.hip", "fcm_column.hip", "fcm_column32.hip", "fcm_noalloc.hip", "fcm_synth.cpp"]
ARCH = os.environ.get("FCM_OFFLOAD_ARCH", "gfx950")

COMMON = ["-O3", "-fPIC", "-std=c++17", "-Wall", "-Wno-unused-result", "-Wno-unused-function"]
HIPFLAGS = [f"--offload-arch={ARCH}", "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
            "-munsafe-fp-atomics"]


def _compile(src):
    path = os.path.join(CSRC, src)
    obj = os.path.join(BUILD, src + ".o")
    if src.endswith(".hip"):
        cmd = [HIPCC, "-x", "hip", *HIPFLAGS, *COMMON, "-c", path, "-o", obj]
    else:
        cmd = [HIPCC, *COMMON, "-c", path, "-o", obj]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"compile failed: {' '.join(cmd)}\n{r.stdout}\n{r.stderr}")
    return obj, r.stderr


def build(jobs=None, verbose=False):
    os.makedirs(BUILD, exist_ok=True)
    jobs = jobs or min(len(SOURCES), os.cpu_count() or 4)
    with cf.ThreadPoolExecutor(max_workers=jobs) as ex:
        results = list(ex.map(_compile, SOURCES))
    if verbose:
        for obj, err in results:
            if err.strip():
                sys.stderr.write(err)
    objs = [o for o, _ in results]
    tmp = LIB + ".tmp"
    cmd = [HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", tmp, *objs]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"link failed: {' '.join(cmd)}\n{r.stdout}\n{r.stderr}")
    os.replace(tmp, LIB)
    build_host()
    return LIB


HOST = os.path.join(HERE, "host")
HOST_LIB = os.path.join(HERE, "libfcm_host.so")
HOST_CHECK = os.path.join(HERE, "fcm_host_check")


def build_host():
    """The reference-named C++ host classes (host/fcm_classes.*) as
    libfcm_host.so over libfcm_hip.so, and the fcm_host_check driver."""
    cxx = os.environ.get("CXX", "g++")
    rpath = "-Wl,-rpath,$ORIGIN"
    lib_cmd = [cxx, "-O2", "-std=c++17", "-fPIC", "-Wall", "-shared", "-o", HOST_LIB + ".tmp",
               os.path.join(HOST, "fcm_classes.cpp"), f"-L{HERE}", "-lfcm_hip", rpath]
    exe_cmd = [cxx, "-O2", "-std=c++17", "-Wall", "-o", HOST_CHECK + ".tmp", os.path.join(HOST, "fcm_host_check.cpp"),
               f"-L{HERE}", "-lfcm_host", "-lfcm_hip", rpath]
    for cmd, out in ((lib_cmd, HOST_LIB), (exe_cmd, HOST_CHECK)):
        r = subprocess.run(cmd, capture_output=True, text=True)
        if r.returncode != 0:
            raise RuntimeError(f"host build failed: {' '.join(cmd)}\n{r.stdout}\n{r.stderr}")
        os.replace(out + ".tmp", out)


if __name__ == "__main__":
    ap = argparse.ArgumentParser()
    ap.add_argument("--jobs", type=int, default=None)
    ap.add_argument("-v", "--verbose", action="store_true")
    a = ap.parse_args()
    print(build(a.jobs, a.verbose))
