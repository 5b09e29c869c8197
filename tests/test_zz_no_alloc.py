"""glx_dev_cg_mixed_noalloc (no reference counterpart): the mixed-precision CG
of glx_dev_cg_mixed with every device buffer taken from the library's fixed
per-GPU scratch array.  Through the C ABI, on two calls in a row -- the
largest size it takes (1024 x 1024) and then a smaller lattice -- its result
must equal glx_dev_cg_mixed's on the same inputs bit for bit (the same
kernels in the same order; only the buffers' addresses and the stream
differ), in both forms of the mixed solver (two iterations per launch and
one), and it must refuse sizes beyond its maximum.
"""
import numpy as np
import pytest

from helpers import lattice_inputs, relerr

pytestmark = pytest.mark.gpu


def _pair(glx, lat, mass, b, link_prec=1, **kw):
    db = glx.DeviceArray(b)
    xa = glx.DeviceArray(np.zeros_like(b))
    xb = glx.DeviceArray(np.zeros_like(b))
    ia = glx.dev_cg_mixed(lat, "staggered_normal", mass, xa, db, link_prec=link_prec, **kw)
    ib = glx.dev_cg_mixed_noalloc(lat, "staggered_normal", mass, xb, db, link_prec=link_prec, **kw)
    return xa.download(), ia, xb.download(), ib


def _same(xa, ia, xb, ib):
    assert (ib.iter, ib.success, ib.ops_count, ib.reliable_updates, ib.name) == \
        (ia.iter, ia.success, ia.ops_count, ia.reliable_updates, ia.name)
    assert ib.resSq == ia.resSq
    assert np.array_equal(xa, xb), relerr(xb, xa)


@pytest.mark.parametrize("sstep", [2, 1])
def test_noalloc_equals_mixed_two_calls(glx, orc, sstep):
    glx.set_sstep(sstep)
    try:
        for (lx, ly, mass, beta) in [(1024, 1024, 0.02, 4.0), (256, 128, 0.05, 6.25)]:
            links, bc, _ = lattice_inputs(orc, lx, ly, beta=beta, seed=900 + lx + ly)
            lat = glx.Lattice(links, lx, ly, f32=True, f16=True)
            xa, ia, xb, ib = _pair(glx, lat, mass, bc, max_iter=100000, eps=1e-10, delta=0.1)
            print(f"{lx}x{ly} s={sstep}: {ib.iter} iterations, {ib.reliable_updates} reliable updates, "
                  f"identical to glx_dev_cg_mixed")
            assert ib.success
            _same(xa, ia, xb, ib)
    finally:
        glx.set_sstep()


def test_noalloc_fp16_links_and_max_iter(glx, orc):
    lx, ly = 256, 64
    links, bc, _ = lattice_inputs(orc, lx, ly, beta=4.0, seed=31)
    lat = glx.Lattice(links, lx, ly, f32=True, f16=True)
    _same(*_pair(glx, lat, 0.05, bc, link_prec=2, max_iter=100000, eps=1e-10, delta=0.1))
    # an unconverged stop at max_iter
    xa, ia, xb, ib = _pair(glx, lat, 0.05, bc, max_iter=9, eps=1e-10, delta=0.1)
    assert not ib.success and ib.iter == 9
    _same(xa, ia, xb, ib)


def test_noalloc_refuses_beyond_maximum(glx, orc):
    links = np.ones(2 * 2048 * 1024, np.complex128)
    lat = glx.Lattice(links, 2048, 1024, f32=True, f16=False)
    b = glx.DeviceArray(nbytes=2048 * 1024 * 16, shape=(2048 * 1024,))
    x = glx.DeviceArray(nbytes=2048 * 1024 * 16, shape=(2048 * 1024,))
    with pytest.raises(glx.GlxError, match="exceed its maximum"):
        glx.dev_cg_mixed_noalloc(lat, "staggered_normal", 0.1, x, b)
    small = glx.Lattice(np.ones(2 * 256 * 64, np.complex128), 256, 64, f32=True, f16=False)
    xs = glx.DeviceArray(np.zeros(256 * 64, np.complex128))
    with pytest.raises(glx.GlxError, match="max_iter beyond"):
        glx.dev_cg_mixed_noalloc(small, "staggered_normal", 0.1, xs, xs, max_iter=1 << 19)
