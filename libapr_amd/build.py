"""Build libaprh.so (HIP, gfx950) in-tree: libapr_amd/_lib/libaprh.so.

    python -m libapr_amd.build          # or: python libapr_amd/build.py

Every .hip under libapr_amd/csrc is compiled with hipcc for gfx950 with FP
contraction disabled (the kernels keep the reference's FP32 operation order)
and packed-FP32 instructions disabled, then linked into one shared library.
libapr_amd/csrc_io/aprfile.cpp (the .apr file format, host code) is linked
into libapr_amd/_lib/libaprh_io.so over the system HDF5 and c-blosc.
"""
import concurrent.futures as cf
import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
OUT_DIR = os.path.join(HERE, "_lib")
OBJ_DIR = os.path.join(HERE, "_lib", "obj")
LIB = os.path.join(OUT_DIR, "libaprh.so")

HIPCC = os.environ.get("HIPCC", shutil.which("hipcc") or "/opt/rocm/bin/hipcc")
ARCH = os.environ.get("APRH_ARCH", "gfx950")

CFLAGS = [
    f"--offload-arch={ARCH}",
    "-O3",
    "-std=c++17",
    "-fPIC",
    "-ffp-contract=off",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-unused-result",
]


def _compile(src):
    obj = os.path.join(OBJ_DIR, os.path.basename(src).replace(".hip", ".o"))
    cmd = [HIPCC, *CFLAGS, "-c", src, "-o", obj]
    r = subprocess.run(cmd, capture_output=True, text=True)
    # the host pass does not know the packed-fp32 feature: that note is harmless
    err = "\n".join(l for l in r.stderr.splitlines() if "packed-fp32-ops" not in l)
    if r.returncode != 0:
        raise RuntimeError(f"hipcc failed for {src}:\n{err}")
    return obj, err


IO_SRC = os.path.join(HERE, "csrc_io", "aprfile.cpp")
IO_LIB = os.path.join(OUT_DIR, "libaprh_io.so")
IO_PREFIX = os.environ.get("APRH_IO_PREFIX", "/opt/conda")   # where hdf5.h / blosc.h and their libraries are


def build_io(verbose=True):
    """libaprh_io.so: the .apr file format over the system HDF5 and c-blosc (host code, g++);
    skipped with a message when those libraries are absent"""
    inc, lib = os.path.join(IO_PREFIX, "include"), os.path.join(IO_PREFIX, "lib")
    if not (os.path.exists(os.path.join(inc, "hdf5.h")) and os.path.exists(os.path.join(inc, "blosc.h"))):
        if verbose:
            print(f"libaprh_io.so not built: hdf5.h / blosc.h not found under {inc}")
        return None
    tmp = IO_LIB + ".tmp"
    cmd = ["g++", "-std=c++17", "-O2", "-fPIC", "-shared", IO_SRC, "-I" + inc, "-L" + OUT_DIR, "-laprh", "-L" + lib,
           "-lhdf5", "-lblosc", "-Wl,-rpath,$ORIGIN", "-Wl,-rpath," + lib, "-o", tmp]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"libaprh_io build failed:\n{r.stderr}")
    os.replace(tmp, IO_LIB)
    if verbose:
        print(f"built {IO_LIB}")
    return IO_LIB


def build(verbose=True):
    os.makedirs(OBJ_DIR, exist_ok=True)
    srcs = sorted(os.path.join(CSRC, f) for f in os.listdir(CSRC) if f.endswith(".hip"))
    jobs = min(len(srcs), int(os.environ.get("MAX_JOBS", "8")))
    with cf.ThreadPoolExecutor(jobs) as ex:
        results = list(ex.map(_compile, srcs))
    for obj, err in results:
        if err.strip() and verbose:
            print(err)
    objs = [o for o, _ in results]
    tmp = LIB + ".tmp"
    cmd = [HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", *objs, "-o", tmp]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"link failed:\n{r.stderr}")
    os.replace(tmp, LIB)
    if verbose:
        print(f"built {LIB}")
    build_io(verbose)
    return LIB


if __name__ == "__main__":
    try:
        build()
    except RuntimeError as e:
        print(e, file=sys.stderr)
        sys.exit(1)
