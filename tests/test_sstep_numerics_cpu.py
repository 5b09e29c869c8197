"""The s-step CG algorithm of the mixed solver (two CG iterations per launch;
no reference counterpart), restated in numpy on the CPU: the same coordinate
algebra as the device prelude -- the Gram matrix of [p, Ap, A^2 p, r, Ar] from
the 12 real sums mu_0..4, Re nu_0..3, rho_0..2, two CG iterations on real
5-vectors, stops on the reliable-update factor / tolerance / max_iter, the
FP64 reliable update and the restart from (p, r_true) -- with FP32 vectors.

Pins the algorithm independently of the GPU: on D^dag D (the oracle's
operator, operators.cpp:444-453) it must converge to the same true residual
as plain CG in FP32 with reliable updates, in the same number of iterations
up to a few, with a solution that matches the FP64 oracle's to the accuracy
the tolerance allows."""
import numpy as np

from helpers import lattice_inputs, relerr


def _ops(orc, links, mass, L):
    def a64(v):
        return orc.apply("staggered_normal", v, links, mass, L, L)

    l32 = links.astype(np.complex64).astype(np.complex128)

    def a32(v):
        return orc.apply("staggered_normal", v.astype(np.complex128), l32, mass, L, L).astype(np.complex64)

    return a64, a32


def _rdot(a, b):
    return float(np.vdot(a.astype(np.complex128), b.astype(np.complex128)).real)


def _gram(Y):
    p, ap, a2p, r, ar = Y
    mu = [_rdot(p, p), _rdot(p, ap), _rdot(ap, ap), _rdot(ap, a2p), _rdot(a2p, a2p)]
    nu = [_rdot(p, r), _rdot(p, ar), _rdot(ap, ar), _rdot(a2p, ar)]
    rho = [_rdot(r, r), _rdot(r, ar), _rdot(ar, ar)]
    G = np.zeros((5, 5))
    for i in range(5):
        for j in range(5):
            pi, pj = i < 3, j < 3
            d = (i if pi else i - 3) + (j if pj else j - 3)
            G[i, j] = mu[d] if (pi and pj) else (rho[d] if not (pi or pj) else nu[d])
    return G


def sstep_cg(a64, a32, b, eps, delta, max_iter):
    """Returns (x, iterations, reliable updates)."""
    B = np.zeros((5, 5))
    B[1, 0] = B[2, 1] = B[4, 3] = 1.0
    x = np.zeros_like(b)
    r64 = b - a64(x)
    r = r64.astype(np.complex64)
    p = np.zeros_like(r)
    xs = np.zeros_like(r)
    tol2 = (eps * np.linalg.norm(b)) ** 2
    maxrr = _rdot(r, r)
    beta_restart, k, reliable, Y = 0.0, 0, 0, None
    while True:
        if Y is None:
            p = (r + np.float32(beta_restart) * p).astype(np.complex64)
        else:
            G = _gram(Y)
            pp, rr, xx = np.eye(5)[0], np.eye(5)[3], np.zeros(5)
            dl = G[3, 3]
            stop = False
            for j in range(2):
                w = B @ pp
                pw = pp @ G @ w
                a = dl / pw if pw != 0 else 0.0
                xx = xx + a * pp
                rr = rr - a * w
                dn = max(rr @ G @ rr, 0.0)
                maxrr = max(maxrr, dn)
                if dn < delta * delta * maxrr or dn < tol2 or dn == 0.0 or k == max_iter - 1:
                    stop = True
                    break
                pp = rr + (dn / dl) * pp
                dl = dn
                k += 1

            def comb(c):
                s = np.zeros_like(Y[0])
                for i in range(5):
                    s = (s + np.float32(c[i]) * Y[i]).astype(np.complex64)
                return s

            xs = (xs + comb(xx)).astype(np.complex64)
            p, r = comb(pp), comb(rr)
            if stop:
                x = x + xs.astype(np.complex128)
                xs[:] = 0
                r64 = b - a64(x)
                rt = _rdot(r64, r64)
                reliable += 1
                if rt == 0.0 or rt < tol2 or k >= max_iter - 1:
                    return x, k + 1, reliable
                r = r64.astype(np.complex64)
                beta_restart, maxrr, k, Y = rt / dl, rt, k + 1, None
                continue
        ap = a32(p)
        Y = [p, ap, a32(ap), r, a32(r)]


def plain_cg(a64, a32, b, eps, delta, max_iter):
    """FP32 CG with the same reliable updates, one iteration at a time."""
    x = np.zeros_like(b)
    r = (b - a64(x)).astype(np.complex64)
    p = np.zeros_like(r)
    xs = np.zeros_like(r)
    tol2 = (eps * np.linalg.norm(b)) ** 2
    rsq = maxrr = _rdot(r, r)
    beta, k, reliable = 0.0, 0, 0
    while True:
        p = (r + np.float32(beta) * p).astype(np.complex64)
        ap = a32(p)
        a = rsq / _rdot(p, ap)
        xs = (xs + np.float32(a) * p).astype(np.complex64)
        r = (r - np.float32(a) * ap).astype(np.complex64)
        rn = _rdot(r, r)
        maxrr = max(maxrr, rn)
        if rn < delta * delta * maxrr or rn < tol2 or k == max_iter - 1:
            x = x + xs.astype(np.complex128)
            xs[:] = 0
            r64 = b - a64(x)
            rt = _rdot(r64, r64)
            reliable += 1
            if rt < tol2 or k >= max_iter - 1:
                return x, k + 1, reliable
            r = r64.astype(np.complex64)
            beta, rsq, maxrr, k = rt / rsq, rt, rt, k + 1
            continue
        beta, rsq, k = rn / rsq, rn, k + 1


def test_sstep_converges_like_cg(orc):
    L, mass = 32, 0.05
    links, b, _ = lattice_inputs(orc, L, L, beta=4.0, seed=3)
    a64, a32 = _ops(orc, links, mass, L)
    xs, ks, rs = sstep_cg(a64, a32, b, 1e-10, 0.1, 5000)
    xp, kp, rp = plain_cg(a64, a32, b, 1e-10, 0.1, 5000)
    r = b - a64(xs)
    assert np.linalg.norm(r) / np.linalg.norm(b) <= 1e-10
    assert abs(ks - kp) <= 3, (ks, kp)
    xo, io = orc.solve("cg", "staggered_normal", b, None, links=links, mass=mass, lx=L, ly=L, max_iter=5000,
                       eps=1e-10)
    cond = (mass * mass + 2.0) / (mass * mass)
    assert relerr(xs, xo) <= 2e-10 * cond
    assert abs(ks - io.iter) <= 0.05 * io.iter + 3, (ks, io.iter)


def test_sstep_max_iter_and_restart(orc):
    """max_iter stops on either iteration of a pair; a large delta forces a
    reliable update (and restart) every few iterations and still converges."""
    L, mass = 16, 0.1
    links, b, _ = lattice_inputs(orc, L, L, beta=4.0, seed=4)
    a64, a32 = _ops(orc, links, mass, L)
    for mi in (1, 4, 5):
        _, k, _ = sstep_cg(a64, a32, b, 1e-10, 0.1, mi)
        assert k == mi
    x, k, rel = sstep_cg(a64, a32, b, 1e-10, 0.6, 5000)
    assert rel > 10
    assert np.linalg.norm(b - a64(x)) / np.linalg.norm(b) <= 1e-10
