"""The wave-march kernel of the plain staggered D and D^dag (glx_dwave.h;
operators.cpp:184-239 and :372-441 are the operators) against the oracle and
against the tiled kernel, on lattices of whole 64-column strips:
  * one strip and several, so that lane 0 and lane 63 take their
    x-neighbours across the periodic boundary and from the next strip;
  * bands of 1 to ly rows, rounds of the march that end past the band's
    last row, ragged last bands, bands marching up and down;
  * complex FP64 (to 1e-14 against the oracle), FP32 with FP32 and FP16
    links (to 1e-6 against the tiled kernel's result with the same links);
  * a lattice whose width is not a multiple of 64 falls back to the tiled
    kernel.
"""
import numpy as np
import pytest

from helpers import lattice_inputs, relerr

pytestmark = pytest.mark.gpu

SHAPES = [(64, 4), (64, 64), (128, 37), (192, 100), (320, 7), (256, 256)]
OPS = ["staggered", "staggered_dagger", "staggered_gamma5", "deo", "doe"]


@pytest.fixture
def dw_reset(glx):
    yield
    glx.set_dwave_rows(-1)


def run(glx, lat, op, m, v, fp32, link_prec, rows):
    glx.set_dwave_rows(rows)
    din, dout = glx.DeviceArray(v), glx.DeviceArray(np.zeros_like(v))
    glx.dev_apply(lat, op, m, dout, din, fp32=fp32, link_prec=link_prec)
    return dout.download()


@pytest.mark.parametrize("shape", SHAPES)
@pytest.mark.parametrize("op", OPS)
def test_dwave_fp64_matches_oracle(glx, orc, dw_reset, shape, op):
    lx, ly = shape
    links, bc, _ = lattice_inputs(orc, lx, ly, beta=4.0, seed=lx + 7 * ly)
    lat = glx.Lattice(links, lx, ly)
    m = 0.3
    want = orc.apply(op, bc, links, m, lx, ly)
    tiled = run(glx, lat, op, m, bc, False, glx.LINKS_F32, 0)
    assert relerr(tiled, want) < 1e-14
    for rows in (-1, 1, 5, 16, 18, ly):
        got = run(glx, lat, op, m, bc, False, glx.LINKS_F32, rows)
        assert glx.get_dwave_rows(lx, ly, 16) > 0
        assert relerr(got, want) < 1e-14, (rows, relerr(got, want))
        assert relerr(got, tiled) < 1e-15, (rows, relerr(got, tiled))


@pytest.mark.parametrize("shape", SHAPES)
@pytest.mark.parametrize("op", OPS)
@pytest.mark.parametrize("f16", [False, True])
def test_dwave_fp32_matches_tiled(glx, orc, dw_reset, shape, op, f16):
    lx, ly = shape
    links, bc, _ = lattice_inputs(orc, lx, ly, beta=4.0, seed=3 * lx + ly)
    lat = glx.Lattice(links, lx, ly, f32=True, f16=f16)
    lp = glx.LINKS_F16 if f16 else glx.LINKS_F32
    m = 0.3
    v = bc.astype(np.complex64)
    want = orc.apply(op, v.astype(np.complex128), links, m, lx, ly)
    tiled = run(glx, lat, op, m, v, True, lp, 0)
    assert relerr(tiled, want) < (2e-3 if f16 else 1e-6)
    for rows in (-1, 4, 7, 16, ly):
        got = run(glx, lat, op, m, v, True, lp, rows)
        assert relerr(got, tiled) < 1e-6, (rows, relerr(got, tiled))
        assert relerr(got, want) < (2e-3 if f16 else 1e-6), (rows, relerr(got, want))


def test_dwave_workload_1024(glx, orc, dw_reset):
    """The 1024^2 lattice of the mixed-CG workload: FP64 against the oracle,
    FP32 against the tiled kernel, both operators."""
    lx = ly = 1024
    links, bc, _ = lattice_inputs(orc, lx, ly, beta=4.0, seed=1024)
    lat = glx.Lattice(links, lx, ly, f32=True)
    v32 = bc.astype(np.complex64)
    for op in OPS:
        want = orc.apply(op, bc, links, 0.02, lx, ly)
        got = run(glx, lat, op, 0.02, bc, False, glx.LINKS_F32, -1)
        assert relerr(got, want) < 1e-14
        tiled = run(glx, lat, op, 0.02, v32, True, glx.LINKS_F32, 0)
        got = run(glx, lat, op, 0.02, v32, True, glx.LINKS_F32, -1)
        assert relerr(got, tiled) < 1e-6


def test_dwave_falls_back_to_tiled(glx, orc, dw_reset):
    lx, ly = 96, 40
    assert glx.get_dwave_rows(lx, ly) == 0
    links, bc, _ = lattice_inputs(orc, lx, ly, beta=4.0, seed=96)
    lat = glx.Lattice(links, lx, ly)
    got = run(glx, lat, "staggered", 0.3, bc, False, glx.LINKS_F32, -1)
    assert relerr(got, orc.apply("staggered", bc, links, 0.3, lx, ly)) < 1e-14


DD_SHAPES = [(64, 4), (64, 64), (128, 37), (190, 100), (320, 7), (256, 256)]


@pytest.fixture
def dd_reset(glx):
    yield
    glx.set_ddwave_rows(-1)


def run_dd(glx, lat, m, v, fp32, link_prec, rows, op="staggered_normal"):
    glx.set_ddwave_rows(rows)
    din, dout = glx.DeviceArray(v), glx.DeviceArray(np.zeros_like(v))
    glx.dev_apply(lat, op, m, dout, din, fp32=fp32, link_prec=link_prec)
    return dout.download()


DD_OPS = ["staggered_normal", "m2mdeodoe"]


@pytest.mark.parametrize("shape", DD_SHAPES)
@pytest.mark.parametrize("op", DD_OPS)
def test_ddwave_fp64_matches_oracle(glx, orc, dd_reset, shape, op):
    """The wave-march D^dag D and m^2 - D_eo D_oe (60 owned columns per
    wave, strips that wrap the lattice, ragged last strips and bands)
    against the oracle."""
    lx, ly = shape
    links, bc, _ = lattice_inputs(orc, lx, ly, beta=4.0, seed=5 * lx + ly)
    lat = glx.Lattice(links, lx, ly)
    m = 0.3
    want = orc.apply(op, bc, links, m, lx, ly)
    for rows in (-1, 1, 3, 10, 17, ly):
        got = run_dd(glx, lat, m, bc, False, glx.LINKS_F32, rows, op)
        assert glx.get_ddwave_rows(lx, ly, 16) > 0
        assert relerr(got, want) < 1e-14, (rows, relerr(got, want))


@pytest.mark.parametrize("shape", DD_SHAPES)
@pytest.mark.parametrize("f16", [False, True])
@pytest.mark.parametrize("op", DD_OPS)
def test_ddwave_fp32_matches_tiled(glx, orc, dd_reset, shape, f16, op):
    lx, ly = shape
    links, bc, _ = lattice_inputs(orc, lx, ly, beta=4.0, seed=7 * lx + ly)
    lat = glx.Lattice(links, lx, ly, f32=True, f16=f16)
    lp = glx.LINKS_F16 if f16 else glx.LINKS_F32
    v = bc.astype(np.complex64)
    ref = run_dd(glx, lat, 0.3, v, True, lp, 0, op)
    for rows in (-1, 2, 10, 17, ly):
        got = run_dd(glx, lat, 0.3, v, True, lp, rows, op)
        assert relerr(got, ref) < 1e-6, (rows, relerr(got, ref))


def test_ddwave_workload_1024(glx, orc, dd_reset):
    """The wave-march D^dag D at its own launch shape on the 1024^2 lattice
    of the mixed-CG workload: FP64 against the oracle, FP32 against the
    row-marching kernel."""
    lx = ly = 1024
    assert glx.get_ddwave_rows(lx, ly, 8) == 10 and glx.get_ddwave_rows(lx, ly, 16) == 10
    links, bc, _ = lattice_inputs(orc, lx, ly, beta=4.0, seed=2024)
    lat = glx.Lattice(links, lx, ly, f32=True)
    want = orc.apply("staggered_normal", bc, links, 0.02, lx, ly)
    got = run_dd(glx, lat, 0.02, bc, False, glx.LINKS_F32, -1)
    assert relerr(got, want) < 1e-14
    v32 = bc.astype(np.complex64)
    ref = run_dd(glx, lat, 0.02, v32, True, glx.LINKS_F32, 0)
    got = run_dd(glx, lat, 0.02, v32, True, glx.LINKS_F32, -1)
    assert relerr(got, ref) < 1e-6
