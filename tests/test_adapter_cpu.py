"""The C++ adapter's host logic (include/qhg_population.hpp) without a GPU:
tests/cpp/adapter_unit.cpp is compiled against the header and the library
and run here."""
import os
import subprocess

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)


def test_adapter_host_logic(tmp_path):
    exe = str(tmp_path / "adapter_unit")
    cmd = ["g++", "-O1", "-std=c++17", "-Wall", "-Werror", "-I" + os.path.join(ROOT, "include"),
           os.path.join(HERE, "cpp", "adapter_unit.cpp"), "-L" + os.path.join(ROOT, "qhg4_amd"), "-lqhg4_amd",
           "-Wl,-rpath," + os.path.join(ROOT, "qhg4_amd"), "-o", exe]
    r = subprocess.run(cmd, capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stdout + r.stderr
    r = subprocess.run([exe], capture_output=True, text=True, timeout=60)
    assert r.returncode == 0 and r.stdout.strip() == "ok", r.stdout + r.stderr
