"""Build libmlhip.so (HIP kernels for gfx950 + C ABI + host C++ drivers) and
libmlio.so (Chombo HDF5 checkpoint/plot files, host only) in-tree.

    python mushy-layer_amd/build.py

Every HIP source is compiled with hipcc for gfx950 only, with floating-point
contraction off (so the FP64 kernels evaluate the reference's expressions as
written) and packed-FP32 instruction generation disabled.  libmlio.so is
plain C++ (g++) against the HDF5 C library of HDF5_DIR (default /opt/conda,
HDF5 1.10); without one it is not built and the package's file functions
raise.
"""
import concurrent.futures as cf
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
LIB = os.path.join(HERE, "libmlhip.so")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")

SOURCES = ["ml_hc_kernels.hip", "ml_transfer_kernels.hip", "ml_fp32r.hip", "ml_vc_kernels.hip", "ml_capi.cpp",
           "ml_hc_engine.cpp", "ml_vc.cpp", "ml_batch.hip", "ml_batch.cpp",
           "ml_adjoint.hip", "ml_noalloc.cpp", "ml_flow_kernels.hip", "ml_flow.cpp", "ml_boxes.hip", "ml_hc_methods.hip",
           "ml_advect.hip", "ml_many.cpp"]

FLAGS = [
    "--offload-arch=gfx950", "-O3", "-fPIC", "-std=c++17", "-ffp-contract=off",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-function", "-Wno-unused-command-line-argument",
    "-I", os.path.join(ROOT, "include"), "-I", CSRC,
]


def _compile(src, obj_dir):
    obj = os.path.join(obj_dir, os.path.basename(src) + ".o")
    cmd = [HIPCC] + FLAGS + ["-x", "hip", "-c", os.path.join(CSRC, src), "-o", obj]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("compile failed: %s\n%s\n%s" % (" ".join(cmd), r.stdout, r.stderr))
    return obj


def build(verbose=False):
    obj_dir = os.path.join(HERE, "_build")
    os.makedirs(obj_dir, exist_ok=True)
    with cf.ThreadPoolExecutor(max_workers=min(8, len(SOURCES))) as ex:
        objs = list(ex.map(lambda s: _compile(s, obj_dir), SOURCES))
    tmp = LIB + ".tmp"
    cmd = [HIPCC, "--offload-arch=gfx950", "-shared", "-fPIC", "-o", tmp] + objs
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("link failed: %s\n%s\n%s" % (" ".join(cmd), r.stdout, r.stderr))
    os.replace(tmp, LIB)
    if verbose:
        print("built", LIB)
    return LIB


IO_LIB = os.path.join(HERE, "libmlio.so")
HDF5_DIR = os.environ.get("HDF5_DIR", "/opt/conda")


def build_io(verbose=False):
    """libmlio.so from io/ml_chombo_h5.cpp; returns its path, or None when no
    HDF5 library is found."""
    inc = os.path.join(HDF5_DIR, "include")
    libdir = os.path.join(HDF5_DIR, "lib")
    if not (os.path.exists(os.path.join(inc, "hdf5.h")) and os.path.exists(os.path.join(libdir, "libhdf5.so"))):
        if verbose:
            print("libmlio.so not built: no HDF5 under %s" % HDF5_DIR)
        return None
    tmp = IO_LIB + ".tmp"
    cmd = ["g++", "-O2", "-fPIC", "-shared", "-std=c++17", "-Wall", "-Wextra", "-I", os.path.join(ROOT, "include"),
           "-I", inc, os.path.join(HERE, "io", "ml_chombo_h5.cpp"), "-o", tmp, "-L", libdir, "-lhdf5",
           "-Wl,-rpath," + libdir]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("libmlio build failed: %s\n%s\n%s" % (" ".join(cmd), r.stdout, r.stderr))
    os.replace(tmp, IO_LIB)
    if verbose:
        print("built", IO_LIB)
    return IO_LIB


if __name__ == "__main__":
    build(verbose=True)
    build_io(verbose=True)
    sys.exit(0)
