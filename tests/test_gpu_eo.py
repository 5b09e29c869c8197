"""Even/odd pipeline on compacted even-site vectors (operators.cpp:543-616
computes it on full-length vectors with the other parity zeroed).

- The compact Schur operator m^2 - D_eo D_oe equals the oracle's m2mdeodoe
  on the even sites (1e-14 in FP64, every launch shape; FP32 to rounding).
- Compact prepare / reconstruct equal the oracle's eoprec_prepare /
  eoprec_reconstruct.
- The compact solve (reference-order CG) takes exactly the oracle's full-
  length eo CG iterations and gives its solution; it solves D x = b.
- Single-pass and mixed-precision solves reach the tolerance.
"""
import numpy as np
import pytest

from helpers import ENVELOPE, eo_envelope, lattice_inputs, relerr

pytestmark = pytest.mark.gpu


def compact(v, lx, ly):
    """Even sites of a full-length vector, in compact order y*(lx/2) + (x>>1)."""
    y, x = np.divmod(np.arange(lx * ly), lx)
    return v[((x + y) % 2) == 0]


def expand(ve, lx, ly):
    out = np.zeros(lx * ly, np.complex128)
    y, x = np.divmod(np.arange(lx * ly), lx)
    out[((x + y) % 2) == 0] = ve
    return out


@pytest.mark.parametrize("lx,ly", [(128, 8), (128, 12), (256, 72), (384, 40)])
@pytest.mark.parametrize("rows", [8, 16, 64])
def test_schur_operator_fp64(glx, orc, rows_reset, lx, ly, rows):
    links, bc, _ = lattice_inputs(orc, lx, ly, beta=2.0, seed=13 + lx + ly)
    m = 0.3
    glx.set_normal_rows(lx, ly, 16, rows)
    ve = compact(bc, lx, ly)
    want = compact(orc.apply("m2mdeodoe", expand(ve, lx, ly), links, m, lx, ly), lx, ly)
    lat = glx.Lattice(links, lx, ly)
    din = glx.DeviceArray(ve)
    dout = glx.DeviceArray(np.zeros_like(ve))
    glx.dev_apply(lat, "staggered_eo_schur", m, dout, din)
    assert relerr(dout.download(), want) < 1e-14


@pytest.mark.parametrize("lx,ly", [(256, 8), (512, 72)])
@pytest.mark.parametrize("link_prec", [1, 2])
def test_schur_operator_fp32(glx, orc, lx, ly, link_prec):
    links, bc, _ = lattice_inputs(orc, lx, ly, beta=2.0, seed=17 + lx)
    m = 0.3
    ve = compact(bc, lx, ly).astype(np.complex64)
    want = compact(orc.apply("m2mdeodoe", expand(ve.astype(np.complex128), lx, ly), links, m, lx, ly), lx, ly)
    lat = glx.Lattice(links, lx, ly)
    din = glx.DeviceArray(ve)
    dout = glx.DeviceArray(np.zeros_like(ve))
    glx.dev_apply(lat, "staggered_eo_schur", m, dout, din, fp32=True, link_prec=link_prec)
    assert relerr(dout.download(), want) < (1e-6 if link_prec == 1 else 2e-3)


def test_prepare_reconstruct(glx, orc):
    lx, ly, m = 128, 12, 0.2
    links, bc, _ = lattice_inputs(orc, lx, ly, beta=3.0, seed=11)
    lat = glx.Lattice(links, lx, ly)
    db = glx.DeviceArray(bc)
    be = glx.DeviceArray(np.zeros(lx * ly // 2, np.complex128))
    glx.dev_eo_prepare(lat, m, be, db)
    assert relerr(be.download(), compact(orc.eoprec_prepare(bc, links, m, lx, ly), lx, ly)) < 1e-14
    xe = orc.Rng(12).gaussian_complex(lx * ly // 2)
    dx = glx.DeviceArray(np.zeros_like(bc))
    glx.dev_eo_reconstruct(lat, m, dx, glx.DeviceArray(xe), db)
    want = orc.eoprec_reconstruct(expand(xe, lx, ly), bc, links, m, lx, ly)
    assert relerr(dx.download(), want) < 1e-14


@pytest.mark.parametrize("lx,ly,m", [(128, 64, 0.1), (256, 128, 0.05)])
def test_eo_solve_matches_oracle_pipeline(glx, orc, lx, ly, m):
    links, bc, _ = lattice_inputs(orc, lx, ly, beta=6.25, seed=21 + lx)
    kw = dict(links=links, mass=m, lx=lx, ly=ly, max_iter=5000, eps=1e-10)
    be = orc.eoprec_prepare(bc, links, m, lx, ly)
    xe_o, io = orc.solve("cg", "m2mdeodoe", be, None, **kw)
    x_o = orc.eoprec_reconstruct(xe_o, bc, links, m, lx, ly)
    lat = glx.Lattice(links, lx, ly)
    dx = glx.DeviceArray(np.zeros_like(bc))
    info, r2 = glx.dev_solve_staggered_eo(lat, m, dx, glx.DeviceArray(bc), "cg", max_iter=5000, eps=1e-10)
    assert (info.iter, info.ops_count, info.success) == (io.iter, io.ops_count, io.success)
    x = dx.download()
    # same iterations; rounding differs (FMA, reduction order) and CG on the
    # Schur operator (kappa ~ (m^2 + 4) / m^2 = 400 .. 1600 here) carries it
    # to ~1e-11; the odd sites (b_o + K x_e) / m carry the even-site error
    # amplified by up to |K| / m <= 2 / m
    assert relerr(compact(x, lx, ly), compact(x_o, lx, ly)) < 1e-10
    assert relerr(x, x_o) < 1e-10 * (1.0 + 2.0 / m)
    r = bc - orc.apply("staggered", x, links, m, lx, ly)
    assert abs(np.sqrt(r2) - np.linalg.norm(r)) <= 1e-6 * np.linalg.norm(r)
    # the odd equations hold by construction and the even residual is the
    # Schur residual / m, with |b_e'| <= (m + 2) |b|:
    # |D x - b| <= 1e-10 (m + 2) / m |b|
    assert np.linalg.norm(r) <= 1e-10 * (m + 2.0) / m * np.linalg.norm(bc) * 1.01


@pytest.mark.parametrize("method", ["cg_fused", "mixed", "mixed_f16"])
def test_eo_solve_fast_methods(glx, orc, method):
    lx, ly, m = 256, 64, 0.05
    links, bc, _ = lattice_inputs(orc, lx, ly, beta=6.25, seed=31)
    lat = glx.Lattice(links, lx, ly)
    dx = glx.DeviceArray(np.zeros_like(bc))
    info, r2 = glx.dev_solve_staggered_eo(lat, m, dx, glx.DeviceArray(bc), method, max_iter=5000, eps=1e-10)
    assert info.success
    x = dx.download()
    xs, err_o = eo_envelope(orc, bc, links, m, lx, ly)
    assert relerr(x, xs) <= ENVELOPE * err_o, (relerr(x, xs), err_o)
    r = bc - orc.apply("staggered", x, links, m, lx, ly)
    assert np.linalg.norm(r) <= 1e-10 * (m + 2.0) / m * np.linalg.norm(bc) * 1.01


def test_eo_rejects_unsupported_size(glx, orc):
    links, bc, _ = lattice_inputs(orc, 24, 10, beta=6.25, seed=3)
    lat = glx.Lattice(links, 24, 10)
    with pytest.raises(glx.GlxError, match="compact even/odd"):
        glx.dev_solve_staggered_eo(lat, 0.1, glx.DeviceArray(np.zeros_like(bc)), glx.DeviceArray(bc))
