"""Builds libluna_hip.so in-tree with hipcc for gfx950.

Each source is compiled to an object in its own hipcc process (they run
concurrently), then the objects are linked into one shared library.
"""
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
SRCS = [os.path.join(HERE, "csrc", n) for n in ("lh_kernels.hip", "lh_fno.hip", "lh_ppt.hip")]
OUT_DIR = os.path.join(HERE, "_build")
OUT = os.path.join(OUT_DIR, "libluna_hip.so")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")

FLAGS = [
    "--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC",
    # packed FP32 instructions stay off: the solver kernels are FP64, and the
    # surrogate's FP32 elementwise code is not worth the risk
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-unused-value", "-Wno-unused-command-line-argument",
]


H5_SRC = os.path.join(HERE, "csrc", "lh_h5.c")
H5_OUT = os.path.join(OUT_DIR, "libluna_h5.so")
HDF5_ROOT = os.environ.get("LUNA_HDF5_ROOT", "/opt/conda")


def build_h5(verbose=False):
    """libluna_h5.so, the HDF5 file layer of HDF5Output (host code, gcc),
    built only where the HDF5 C library and its header are found.  Returns its
    path, or None when HDF5 is absent (HDF5Output then raises on use)."""
    inc, lib = os.path.join(HDF5_ROOT, "include"), os.path.join(HDF5_ROOT, "lib")
    if not (os.path.isfile(os.path.join(inc, "hdf5.h")) and os.path.isfile(os.path.join(lib, "libhdf5.so"))):
        if verbose:
            sys.stderr.write(f"HDF5 not found under {HDF5_ROOT}: HDF5Output not built\n")
        return None
    os.makedirs(OUT_DIR, exist_ok=True)
    tmp = H5_OUT + f".tmp{os.getpid()}"
    res = subprocess.run(["gcc", "-O2", "-std=gnu99", "-fPIC", "-shared", "-Wall", "-I" + inc, H5_SRC,
                          "-L" + lib, "-lhdf5", "-Wl,-rpath," + lib, "-o", tmp], capture_output=True, text=True)
    if res.returncode != 0:
        sys.stderr.write(res.stdout + res.stderr)
        raise RuntimeError("building libluna_h5.so failed")
    os.replace(tmp, H5_OUT)
    return H5_OUT


def build(verbose=False):
    os.makedirs(OUT_DIR, exist_ok=True)
    tag = f".tmp{os.getpid()}"
    objs = [os.path.join(OUT_DIR, os.path.basename(s).replace(".hip", ".o") + tag) for s in SRCS]
    procs = [subprocess.Popen([HIPCC] + FLAGS + ["-c", "-o", o, s], stdout=subprocess.PIPE,
                              stderr=subprocess.PIPE, text=True) for s, o in zip(SRCS, objs)]
    failed = []
    for p, s in zip(procs, SRCS):
        out, err = p.communicate()
        if p.returncode != 0:
            sys.stderr.write(out + err)
            failed.append(s)
        elif verbose:
            sys.stderr.write(err)
    try:
        if failed:
            raise RuntimeError("hipcc failed for " + " ".join(failed))
        tmp = OUT + tag
        res = subprocess.run([HIPCC, "--offload-arch=gfx950", "-shared", "-fPIC", "-o", tmp] + objs,
                             capture_output=True, text=True)
        if res.returncode != 0:
            sys.stderr.write(res.stdout + res.stderr)
            raise RuntimeError("linking libluna_hip.so failed")
        os.replace(tmp, OUT)
        build_h5(verbose)
    finally:
        for o in objs:
            if os.path.exists(o):
                os.remove(o)
    return OUT


if __name__ == "__main__":
    print(build(verbose=True))
