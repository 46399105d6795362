"""Build the C++ host driver over include/qhg_population.hpp:
tests/cpp/_build/popbase_driver (linked to qhg4_amd/libqhg4_amd.so)."""
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


if __name__ == "__main__":
    print(build(verbose=True))
