"""The wide launch shapes large lattices take by default (128 and 256 rows per
block), forced on a small lattice: 128 x 520 is two full bands of 256 rows
(the second marching upwards in the s-step kernel) and a ragged band of 8;
at 128 rows, four full bands and the ragged one.  Each kernel against the
oracle (operators.cpp:444-453 and :566-588 are the operators):
  * the row-marching D^dag D, FP64 (to 1e-14) and FP32 (against the tiled
    kernel);
  * the compact even/odd Schur operator m^2 - D_eo D_oe, FP64;
  * the single-pass FP64 CG and the mixed CG in s-step form, and the mixed
    even/odd solve, with the true residual recomputed by the oracle's FP64
    operator.
"""
import numpy as np
import pytest

from helpers import lattice_inputs, relerr

pytestmark = pytest.mark.gpu

LX, LY = 128, 520
WIDE = [128, 256]


def compact(v, lx, ly):
    y, x = np.divmod(np.arange(lx * ly), lx)
    return v[((x + y) % 2) == 0]


def expand(ve, lx, ly):
    out = np.zeros(lx * ly, np.complex128)
    y, x = np.divmod(np.arange(lx * ly), lx)
    out[((x + y) % 2) == 0] = ve
    return out


@pytest.fixture(scope="module")
def inputs(orc):
    links, bc, _ = lattice_inputs(orc, LX, LY, beta=4.0, seed=520)
    return links, bc


@pytest.mark.parametrize("rows", WIDE)
def test_wide_normal_fp64(glx, orc, rows_reset, inputs, rows):
    links, bc = inputs
    glx.set_normal_rows(LX, LY, 16, rows)
    assert glx.get_normal_rows(LX, LY, 16) == rows
    want = orc.apply("staggered_normal", bc, links, 0.3, LX, LY)
    got = glx.apply("staggered_normal", bc, links, 0.3, LX, LY)
    assert relerr(got, want) < 1e-14, relerr(got, want)


@pytest.mark.parametrize("rows", WIDE)
def test_wide_normal_fp32(glx, orc, rows_reset, inputs, rows):
    links, bc = inputs
    lat = glx.Lattice(links, LX, LY)
    b32 = bc.astype(np.complex64)
    din, dout = glx.DeviceArray(b32), glx.DeviceArray(np.zeros_like(b32))
    glx.set_normal_rows(LX, LY, 8, 0)
    glx.dev_apply(lat, "staggered_normal", 0.2, dout, din, fp32=True)
    tiled = dout.download()
    glx.set_normal_rows(LX, LY, 8, rows)
    glx.dev_apply(lat, "staggered_normal", 0.2, dout, din, fp32=True)
    assert relerr(dout.download(), tiled) < 1e-6
    want = orc.apply("staggered_normal", b32.astype(np.complex128), links, 0.2, LX, LY)
    assert relerr(dout.download(), want) < 1e-6


@pytest.mark.parametrize("rows", WIDE)
def test_wide_eo_schur_fp64(glx, orc, rows_reset, inputs, rows):
    links, bc = inputs
    m = 0.3
    glx.set_normal_rows(LX, LY, 16, rows)
    ve = compact(bc, LX, LY)
    lat = glx.Lattice(links, LX, LY)
    din, dout = glx.DeviceArray(ve), glx.DeviceArray(np.zeros_like(ve))
    glx.dev_apply(lat, "staggered_eo_schur", m, dout, din)
    want = compact(orc.apply("m2mdeodoe", expand(ve, LX, LY), links, m, LX, LY), LX, LY)
    assert relerr(dout.download(), want) < 1e-14


@pytest.mark.parametrize("rows", WIDE)
def test_wide_single_pass_cg(glx, orc, rows_reset, inputs, rows):
    links, bc = inputs
    m = 0.05
    glx.set_normal_rows(LX, LY, 16, rows)
    lat = glx.Lattice(links, LX, LY)
    dx = glx.DeviceArray(np.zeros_like(bc))
    info = glx.dev_cg_fused(lat, "staggered_normal", m, dx, glx.DeviceArray(bc), max_iter=20000, eps=1e-10)
    x = dx.download()
    r = bc - orc.apply("staggered_normal", x, links, m, LX, LY)
    assert info.success and np.linalg.norm(r) / np.linalg.norm(bc) <= 1.001e-10
    xo, io = orc.solve("cg", "staggered_normal", bc, None, links=links, mass=m, lx=LX, ly=LY, max_iter=20000,
                       eps=1e-10)
    assert abs(info.iter - io.iter) <= 2 and relerr(x, xo) < 1e-11


@pytest.mark.parametrize("rows", WIDE)
@pytest.mark.parametrize("link_prec", [1, 2])
def test_wide_sstep_mixed(glx, orc, inputs, rows, link_prec):
    links, bc = inputs
    m = 0.05
    lat = glx.Lattice(links, LX, LY)
    glx.set_sstep(2, rows)
    try:
        dx = glx.DeviceArray(np.zeros_like(bc))
        info = glx.dev_cg_mixed(lat, "staggered_normal", m, dx, glx.DeviceArray(bc), max_iter=20000, eps=1e-10,
                                delta=0.1, link_prec=link_prec)
    finally:
        glx.set_sstep()
    x = dx.download()
    r = bc - orc.apply("staggered_normal", x, links, m, LX, LY)
    relres = np.linalg.norm(r) / np.linalg.norm(bc)
    assert info.success and relres <= 1.001e-10, relres
    assert abs(np.sqrt(info.resSq) / np.linalg.norm(bc) - relres) <= 1e-3 * relres
    xo, io = orc.solve("cg", "staggered_normal", bc, None, links=links, mass=m, lx=LX, ly=LY, max_iter=20000,
                       eps=1e-10)
    assert relerr(x, xo) < 5e-11 and info.iter <= 1.2 * io.iter + 10


@pytest.mark.parametrize("rows", WIDE)
def test_wide_eo_mixed_solve(glx, orc, inputs, rows):
    links, bc = inputs
    m = 0.05
    lat = glx.Lattice(links, LX, LY)
    glx.set_sstep(2, rows)
    try:
        dx = glx.DeviceArray(np.zeros_like(bc))
        info, r2 = glx.dev_solve_staggered_eo(lat, m, dx, glx.DeviceArray(bc), "mixed", max_iter=20000, eps=1e-10)
    finally:
        glx.set_sstep()
    x = dx.download()
    r = bc - orc.apply("staggered", x, links, m, LX, LY)
    assert info.success
    assert abs(np.sqrt(r2) - np.linalg.norm(r)) <= 1e-6 * np.linalg.norm(r)
    # even residual = Schur residual / m with |b_e'| <= (m + 2) |b|
    assert np.linalg.norm(r) <= 1e-10 * (m + 2.0) / m * np.linalg.norm(bc) * 1.01
