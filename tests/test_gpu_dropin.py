"""The reference-style C++ driver (tests/dropin/square_laplace_driver.cpp),
compiled against the compat headers and linked with libglx.so, runs on the
GPU: registered operators, a restarted solver, multishift CG with the shifts
array restored, and the driver's own host callback all converge."""
import os
import re
import subprocess

import pytest

pytestmark = pytest.mark.gpu
ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_dropin_driver_runs(glx):
    exe = os.path.join(ROOT, "tests", "dropin", "square_laplace_driver")
    if not os.path.exists(exe):
        subprocess.run(["make", "-C", os.path.dirname(exe)], check=True)
    out = subprocess.run([exe], capture_output=True, text=True, timeout=120)
    assert out.returncode == 0, out.stdout + out.stderr
    t = out.stdout
    assert re.search(r"\[CG\]: CG Success Y Iter \d+ Ops \d+ RelRes", t)
    m = re.search(r"\[TEST\]: cg CG iter (\d+) ops (\d+) success 1 relres (\S+)", t)
    assert m and int(m.group(2)) == int(m.group(1)) + 2 and float(m.group(3)) < 1e-10
    assert "[TEST]: bicgstab" in t and re.search(r"bicgstab iter \d+ success 1", t)
    assert "[TEST]: gcr_restart GCR(32) success 1" in t
    m = re.search(r"cg_m iter \d+ resSq (\S+) shifts (\S+) (\S+) resSqmrhs0 (\S+)", t)
    assert m and float(m.group(1)) == 0.0 and float(m.group(2)) == 0.01 and float(m.group(3)) == 0.005
    m = re.search(r"host_callback iter \d+ success 1 relres (\S+)", t)
    assert m and float(m.group(1)) < 1e-10
    # minv_unpreconditioned (generic_inverter.cpp:18-188)
    assert "[TEST]: dispatch cg CG success 1" in t
    assert re.search(r"\[TEST\]: dispatch bicgstab_restart BiCGStab\(64\) success [01]", t)
    assert "[TEST]: dispatch invalid iter 0" in t
    assert "[TEST]: dispatch cr refused" in t
