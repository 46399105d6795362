"""Build the C++ test drivers:
- tests/cpp/_build/popbase_driver: the host driver over
  include/qhg_population.hpp (linked to qhg4_amd/libqhg4_amd.so);
- tests/cpp/_build/sanitize_driver (build_sanitize): the oracle, config.cpp
  and qdf_io.cpp compiled with -fsanitize=address,undefined, host only (no
  HIP source is compiled);
- tests/cpp/_build/dispatch_probe (build_probe): the kernel-boundary cost
  probe, run under rocprofv3 for the step kernels' design."""
import os
import subprocess

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
OUT = os.path.join(HERE, "_build", "popbase_driver")


def build(verbose=False):
    os.makedirs(os.path.dirname(OUT), exist_ok=True)
    cmd = ["g++", "-O2", "-std=c++17", "-Wall", "-I" + os.path.join(ROOT, "include"),
           os.path.join(HERE, "popbase_driver.cpp"), "-L" + os.path.join(ROOT, "qhg4_amd"), "-lqhg4_amd",
           "-Wl,-rpath,$ORIGIN/../../../qhg4_amd", "-o", OUT + ".tmp"]
    if verbose:
        print(" ".join(cmd), flush=True)
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("g++ (popbase_driver) failed:\n" + r.stdout + r.stderr)
    os.replace(OUT + ".tmp", OUT)
    return OUT


SAN_OUT = os.path.join(HERE, "_build", "sanitize_driver")
HDF5 = os.environ.get("QHG_HDF5_PREFIX", "/opt/conda")
SAN = ["-fsanitize=address,undefined", "-fno-sanitize-recover=all", "-fno-omit-frame-pointer", "-g", "-O1"]


def build_sanitize(verbose=False):
    """The sanitizer build (needs the HDF5 headers and the library link of
    qhg4_amd/_hdf5, which the package build makes)."""
    os.makedirs(os.path.dirname(SAN_OUT), exist_ok=True)
    obj = os.path.join(HERE, "_build", "qhg_oracle_san.o")
    cc = ["gcc", "-std=c11", "-fopenmp", "-ffp-contract=off", "-fno-fast-math"] + SAN + [
        "-c", os.path.join(ROOT, "oracle", "qhg_oracle.c"), "-o", obj]
    hdf5_dir = os.path.join(ROOT, "qhg4_amd", "_hdf5")
    cxx = ["g++", "-std=c++17", "-fopenmp", "-Wall"] + SAN + [
        "-I" + os.path.join(ROOT, "include"), "-I" + os.path.join(ROOT, "oracle"), "-I" + os.path.join(HDF5, "include"),
        os.path.join(HERE, "sanitize_driver.cpp"), os.path.join(ROOT, "qhg4_amd", "csrc", "config.cpp"),
        os.path.join(ROOT, "qhg4_amd", "csrc", "qdf_io.cpp"), obj, "-L" + hdf5_dir, "-l:libhdf5.so.103",
        "-Wl,-rpath," + hdf5_dir, "-lm", "-o", SAN_OUT + ".tmp"]
    for cmd in (cc, cxx):
        if verbose:
            print(" ".join(cmd), flush=True)
        r = subprocess.run(cmd, capture_output=True, text=True)
        if r.returncode != 0:
            raise RuntimeError("sanitizer build failed:\n" + r.stdout + r.stderr)
    os.replace(SAN_OUT + ".tmp", SAN_OUT)
    return SAN_OUT


PROBE_OUT = os.path.join(HERE, "_build", "dispatch_probe")


def build_probe(verbose=False):
    """The dispatch-cost probe (tests/cpp/dispatch_probe.hip), for gfx950."""
    os.makedirs(os.path.dirname(PROBE_OUT), exist_ok=True)
    cmd = ["/opt/rocm/bin/hipcc", "--offload-arch=gfx950", "-O3", "-Xclang", "-target-feature", "-Xclang",
           "-packed-fp32-ops", os.path.join(HERE, "dispatch_probe.hip"), "-o", PROBE_OUT + ".tmp"]
    if verbose:
        print(" ".join(cmd), flush=True)
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("hipcc (dispatch_probe) failed:\n" + r.stdout + r.stderr)
    os.replace(PROBE_OUT + ".tmp", PROBE_OUT)
    return PROBE_OUT


if __name__ == "__main__":
    print(build(verbose=True))
