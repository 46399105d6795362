"""Build the HIP library in-tree: qhg4_amd/libqhg4_amd.so (gfx950).

    python -m qhg4_amd.build
"""
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
LIB = os.path.join(HERE, "libqhg4_amd.so")
SOURCES = ["qhg_gpu.hip", "config.cpp"]
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")

FLAGS = [
    "--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC", "-shared",
    # bit-exact FP64 arithmetic: no contraction into FMA
    "-ffp-contract=off",
    # no packed FP32 instructions
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall",
]


# QDF (HDF5) I/O: a host library of its own, linked to the image's HDF5.
# The HDF5 prefix also holds an older libstdc++ than the ROCm libraries
# need, so the library's search path is a private directory holding only a
# link to libhdf5 (qhg4_amd/_hdf5), not the prefix's lib directory.
QDF_LIB = os.path.join(HERE, "libqhg4_qdf.so")
HDF5 = os.environ.get("QHG_HDF5_PREFIX", "/opt/conda")
HDF5_SONAME = "libhdf5.so.103"
HDF5_LINK_DIR = os.path.join(HERE, "_hdf5")
QDF_CMD = ["g++", "-O2", "-std=c++17", "-fPIC", "-shared", "-Wall", "-I" + os.path.join(HDF5, "include"),
           os.path.join(CSRC, "qdf_io.cpp"), "-L" + HDF5_LINK_DIR, "-l:" + HDF5_SONAME, "-L" + HERE,
           "-lqhg4_amd", "-Wl,-rpath,$ORIGIN/_hdf5", "-Wl,-rpath,$ORIGIN"]


# PopBase plugins (qhg4_amd/plugin): they compile against a QHG4 source
# tree's headers and are built where one is present.
QHG4_TREE = os.environ.get("QHG4_TREE", "/root/reference/QHG4")
PLUGIN_DIR = os.path.join(HERE, "plugin")


def _hdf5_link():
    os.makedirs(HDF5_LINK_DIR, exist_ok=True)
    link = os.path.join(HDF5_LINK_DIR, HDF5_SONAME)
    target = os.path.join(HDF5, "lib", HDF5_SONAME)
    if os.path.islink(link) and os.readlink(link) == target:
        return
    if os.path.lexists(link):
        os.remove(link)
    os.symlink(target, link)


def _run(cmd, what, verbose):
    if verbose:
        print(" ".join(cmd), flush=True)
    r = subprocess.run(cmd, capture_output=True, text=True)
    msgs = [ln for ln in (r.stdout + r.stderr).splitlines() if "packed-fp32-ops" not in ln]
    if r.returncode != 0:
        sys.stderr.write("\n".join(msgs) + "\n")
        raise RuntimeError("%s failed (%d)" % (what, r.returncode))


def build(verbose=False):
    _run([HIPCC] + FLAGS + ["-o", LIB + ".tmp"] + [os.path.join(CSRC, s) for s in SOURCES], "hipcc", verbose)
    os.replace(LIB + ".tmp", LIB)
    if os.path.exists(os.path.join(HDF5, "include", "hdf5.h")):
        _hdf5_link()
        _run(QDF_CMD + ["-o", QDF_LIB + ".tmp"], "g++ (QDF library)", verbose)
        os.replace(QDF_LIB + ".tmp", QDF_LIB)
        if os.path.isdir(os.path.join(QHG4_TREE, "core")):
            _run(["make", "-s", "-B", "-j8", "-C", PLUGIN_DIR, "QHG4=" + QHG4_TREE], "make (PopBase plugins)", verbose)
    return LIB


if __name__ == "__main__":
    print(build(verbose=True))
