"""The PopBase plugin (qhg4_amd/plugin) against a QHG4 source tree: it
compiles with the reference's own headers (PopBase, SCellGrid,
ParamProvider2, ...), and the linked Wrapper library exports exactly the
entry points DynPopFactory looks up (getInfo, createPop;
QHG4/populations/DynPopFactory.cpp:15-18), needing from the executable only
QHG4's own symbols.  Skipped where no QHG4 tree is present."""
import os
import shutil
import subprocess

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
PLUGIN = os.path.join(os.path.dirname(HERE), "qhg4_amd", "plugin")
QHG4 = "/root/reference/QHG4"

pytestmark = pytest.mark.skipif(not os.path.isdir(os.path.join(QHG4, "core")), reason="no QHG4 source tree")


def test_plugin_compiles_against_the_reference_headers():
    r = subprocess.run(["make", "-s", "check", "QHG4=" + QHG4], cwd=PLUGIN, capture_output=True, text=True,
                       timeout=300)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "warning" not in (r.stdout + r.stderr)


def test_wrapper_exports(tmp_path):
    for f in ("GpuPop.cpp", "GpuPop.h", "GpuPopWrapper.cpp", "Makefile"):
        shutil.copy(os.path.join(PLUGIN, f), str(tmp_path))
    engine = os.path.dirname(PLUGIN)
    r = subprocess.run(["make", "-s", "GpuOoAPopWrapper.so", "QHG4=" + QHG4, "ENGINE=" + engine],
                       cwd=str(tmp_path), capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stdout + r.stderr
    so = str(tmp_path / "GpuOoAPopWrapper.so")
    defined = subprocess.run(["nm", "-D", "--defined-only", so], capture_output=True, text=True).stdout
    names = {ln.split()[-1] for ln in defined.splitlines() if " T " in ln}
    assert {"getInfo", "createPop"} <= names
    undefined = subprocess.run(["nm", "-DC", "--undefined-only", so], capture_output=True, text=True).stdout
    qhg = [ln.split(None, 1)[1] for ln in undefined.splitlines()
           if ln.strip().startswith("U") and ("ArrayShare" in ln or "Observable" in ln or "SCellGrid" in ln)]
    assert sorted(qhg) == ["ArrayShare::setInstance(ArrayShare*)", "Observable::Observable()"]
