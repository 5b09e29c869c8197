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


def test_stencil_driver_runs(glx):
    """tests/dropin/stencil_driver.cpp: Lattice, stencil_2d, the stencil
    builders, eo-prec helpers, generate_stencil_2d and the nc = 2 pieces
    through the compat headers (coarse_stencil.h:14-173,
    operators_stencil.h:1-40), solves by pointer identity on the GPU."""
    exe = os.path.join(ROOT, "tests", "dropin", "stencil_driver")
    if not os.path.exists(exe):
        subprocess.run(["make", "-C", os.path.dirname(exe), "stencil_driver"], check=True)
    out = subprocess.run([exe], capture_output=True, text=True, timeout=120)
    assert out.returncode == 0, out.stdout + out.stderr
    t = out.stdout
    val = lambda pat: float(re.search(pat, t).group(1))
    assert val(r"stencil_vs_function (\S+)") < 1e-15
    # even-site relres 1e-10 bounds the full |D x - b| / |b| by 1e-10 (m + 2) / m
    m = re.search(r"eo_solve success (\d) iter (\d+) residual (\S+)", t)
    assert m and m.group(1) == "1" and float(m.group(3)) < 1.01e-10 * (0.05 + 2) / 0.05
    assert re.search(r"\[EO\]: CG Success Y", t)
    assert val(r"generated_normal (\S+)") < 1e-14 and "generated 1" in t
    assert val(r"directions_sum (\S+)") < 1e-14
    assert val(r"nc2 tb_bottom_rows (\S+)") == 0.0
    m = re.search(r"nc2_bicgstab success (\d) residual (\S+)", t)
    assert m and m.group(1) == "1" and float(m.group(2)) < 1.01e-10
