"""Gauge fields generated and measured in HBM (glx_lattice_create_gaussian,
glx_lattice_gaussian_links, glx_lattice_get_links, glx_lattice_plaquette)
against the oracle: the same Philox4x32-10 draws on the CPU, and
get_plaquette_u1 (u1_utils.cpp:190-207) on the same links.

Tolerances: the device and the CPU evaluate log, sin and cos with different
(faithfully rounded) libraries, so links agree to a few ulp of the angle,
bounded here by 1e-13; the plaquette is a reordered sum of V unit-modulus
terms, bounded by 1e-13 on the mean."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("lx,ly,beta,seed,stream", [(64, 48, 4.0, 1337, 0), (30, 10, 6.25, 2**40 + 5, 7),
                                                    (16, 16, 0.0, 99, 1), (40, 8, -2.0, 3, 0)])
def test_device_field_matches_oracle(glx, orc, lx, ly, beta, seed, stream):
    print(f"case lx={lx} ly={ly} beta={beta} seed={seed} stream={stream}", flush=True)
    lat = glx.Lattice.gaussian(lx, ly, beta, seed, stream)
    got = lat.get_links()
    want = orc.philox_gauge(lx, ly, beta, seed, stream)
    assert np.max(np.abs(got - want)) < 1e-13
    p = lat.plaquette()
    assert abs(p - orc.plaquette(got, lx, ly)) < 1e-13
    lat.close()


def test_plaquette_of_host_fields(glx, orc):
    g = orc.Rng(1337)
    for lx, ly, beta in ((16, 16, 3.0), (128, 96, 10.0), (2, 2, 1.0)):
        u = g.gauss_gauge(lx, ly, beta)
        lat = glx.Lattice(u, lx, ly)
        assert abs(lat.plaquette() - orc.plaquette(u, lx, ly)) < 1e-13
        assert np.array_equal(lat.get_links(), u)
        lat.close()
    lat = glx.Lattice(orc.unit_gauge(32, 32), 32, 32)
    assert lat.plaquette() == 1.0 + 0j
    lat.close()


def test_plaquette_gauge_invariant(glx, orc):
    g = orc.Rng(7)
    lx, ly = 64, 32
    u = g.gauss_gauge(lx, ly, 2.0)
    ug = orc.apply_gauge_trans(u, g.rand_trans(lx, ly), lx, ly)
    a, b = glx.Lattice(u, lx, ly), glx.Lattice(ug, lx, ly)
    assert abs(a.plaquette() - b.plaquette()) < 1e-13
    a.close()
    b.close()


def test_regenerate_refreshes_every_copy(glx, orc):
    """Regenerated links replace the FP64 field and its FP32 / FP16 copies:
    the operator applied with each link precision agrees with the oracle's
    operator on the new field to that precision."""
    lx, ly, m = 64, 64, 0.1
    lat = glx.Lattice.gaussian(lx, ly, 4.0, seed=1)
    lat.regenerate(4.0, seed=2, stream=5)
    u = lat.get_links()
    assert np.array_equal(u, glx.Lattice.gaussian(lx, ly, 4.0, seed=2, stream=5).get_links())
    rng = np.random.default_rng(0)
    v = rng.standard_normal(lx * ly) + 1j * rng.standard_normal(lx * ly)
    want = orc.apply("staggered", v, u, m, lx, ly)
    inp, out = glx.DeviceArray(v), glx.DeviceArray(nbytes=v.nbytes)
    glx.dev_apply(lat, "staggered", m, out, inp)
    assert np.max(np.abs(out.download() - want)) < 1e-13
    v32 = v.astype(np.complex64)
    want = orc.apply("staggered", v32.astype(np.complex128), u, m, lx, ly)
    in32, out32 = glx.DeviceArray(v32), glx.DeviceArray(nbytes=v32.nbytes, dtype=np.complex64)
    for prec, tol in ((glx.LINKS_F32, 1e-5), (glx.LINKS_F16, 1e-2)):
        glx.dev_apply(lat, "staggered", m, out32, in32, fp32=True, link_prec=prec)
        err = np.max(np.abs(out32.download() - want))
        assert err < tol, (prec, err)
    lat.close()


def test_solve_on_device_field(glx, orc):
    """A mixed-precision solve on a field that never left HBM meets the FP64
    true-residual target, checked with the oracle's operator."""
    lx = ly = 128
    m = 0.05
    lat = glx.Lattice.gaussian(lx, ly, 6.25, seed=1337)
    u = lat.get_links()
    rng = np.random.default_rng(1)
    b = rng.standard_normal(lx * ly) + 1j * rng.standard_normal(lx * ly)
    x, bd = glx.DeviceArray(np.zeros_like(b)), glx.DeviceArray(b)
    info = glx.dev_cg_mixed(lat, "staggered_normal", m, x, bd, eps=1e-10)
    assert info.success
    r = b - orc.apply("staggered_normal", x.download(), u, m, lx, ly)
    assert np.linalg.norm(r) <= 1.05e-10 * np.linalg.norm(b)
    lat.close()
