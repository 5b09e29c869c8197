"""The reference's own drop-in consumer compiles unchanged against the compat
headers and links libglx.so (no GPU needed to build).

square_laplace.cpp includes generic_inverters.h, generic_inverters_precond.h
and generic_vector.h and calls minv_vector_gcr with its own host callback.
It is copied from the reference tree (read where it lies; nothing of it is
kept in this repository) into a temporary directory, compiled with g++ with
only the compat directory on the include path, and linked with libglx.so.
Skipped where the reference tree is absent.
"""
import os
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
REF = "/root/reference"
COMPAT = os.path.join(ROOT, "generic-linalg_amd", "compat")
LIBDIR = os.path.join(ROOT, "generic-linalg_amd")


@pytest.mark.skipif(not os.path.exists(os.path.join(REF, "square_laplace.cpp")), reason="reference tree absent")
@pytest.mark.skipif(not os.path.exists(os.path.join(LIBDIR, "libglx.so")), reason="libglx.so not built")
def test_square_laplace_builds_unchanged(tmp_path):
    src = tmp_path / "square_laplace.cpp"
    shutil.copyfile(os.path.join(REF, "square_laplace.cpp"), src)
    exe = tmp_path / "square_laplace"
    cmd = ["g++", "-O2", "-std=c++11", "-I", COMPAT, str(src), "-L", LIBDIR, "-lglx", f"-Wl,-rpath,{LIBDIR}",
           "-o", str(exe)]
    out = subprocess.run(cmd, capture_output=True, text=True, timeout=300)
    assert out.returncode == 0, out.stderr
    assert exe.exists()
    # every undefined symbol of the program resolves in libglx.so or the C/C++ runtime
    nm = subprocess.run(["nm", "-D", "--undefined-only", str(exe)], capture_output=True, text=True, check=True)
    undef = {l.split()[-1] for l in nm.stdout.splitlines() if l.strip()}
    lib = subprocess.run(["nm", "-D", "--defined-only", os.path.join(LIBDIR, "libglx.so")], capture_output=True,
                         text=True, check=True)
    defined = {l.split()[-1] for l in lib.stdout.splitlines() if l.strip()}
    solver = [s for s in undef if "minv_vector" in s]
    assert solver and all(s in defined for s in solver), solver


def test_precond_header_interface():
    """The compat generic_inverters_precond.h keeps the reference's enum
    values and the params fields (generic_inverters_precond.h:31-51)."""
    text = open(os.path.join(COMPAT, "generic_inverters_precond.h")).read()
    for item in ("MINV_PRE_CG = 0", "MINV_PRE_FPCG = 1", "MINV_PRE_VPGCR = 2", "MINV_PRE_BICGSTAB = 3",
                 "MINV_PRE_INVALID = -1", "double tol;", "int max_iters;", "bool restart;", "int restart_freq;"):
        assert item in text, item


PRECOND_PROBE = r"""
#include <cstdio>
#include <stdexcept>
#include "generic_inverters_precond.h"
static void mv(double* l, double* r, void*) { l[0] = r[0]; }
int main() {
  double x = 0.0, b = 1.0;
  minv_inverter_precond_params p = {1e-10, 10, false, 0};
  inversion_info a = minv_preconditioned(&x, &b, 1, MINV_PRE_INVALID, p, mv, 0, 0, 0);
  std::printf("invalid iter %d success %d\n", a.iter, int(a.success));
  try {
    minv_preconditioned(&x, &b, 1, MINV_PRE_CG, p, mv, 0, 0, 0);
    std::printf("cg ran\n");
  } catch (const std::runtime_error& e) {
    std::printf("cg refused: %s\n", e.what());
  }
  return 0;
}
"""


@pytest.mark.skipif(not os.path.exists(os.path.join(LIBDIR, "libglx.so")), reason="libglx.so not built")
def test_minv_preconditioned_refuses(tmp_path):
    """minv_preconditioned returns a default inversion_info for
    MINV_PRE_INVALID and throws for the solvers this library does not carry
    (no device work on either path)."""
    src = tmp_path / "probe.cpp"
    src.write_text(PRECOND_PROBE)
    exe = tmp_path / "probe"
    subprocess.run(["g++", "-O1", "-std=c++11", "-I", COMPAT, str(src), "-L", LIBDIR, "-lglx",
                    f"-Wl,-rpath,{LIBDIR}", "-o", str(exe)], check=True, timeout=300)
    out = subprocess.run([str(exe)], capture_output=True, text=True, timeout=60)
    assert out.returncode == 0, out.stderr
    assert "invalid iter 0 success 0" in out.stdout
    assert "cg refused: minv_preconditioned" in out.stdout
