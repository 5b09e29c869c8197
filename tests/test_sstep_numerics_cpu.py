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

from helpers import ENVELOPE, cg_envelope, lattice_inputs, relerr


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
    x_star, err_o, xo, io = cg_envelope(orc, "staggered_normal", b, links, mass, L, L)
    # numpy's FP32 arithmetic (no fused multiply-adds) takes another
    # trajectory than the GPU's: 7.6 times the FP64 CG's error at the same
    # residual here (both stop near 8e-11)
    assert relerr(xs, x_star) <= 2.5 * ENVELOPE * err_o, (relerr(xs, x_star), err_o)
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


def _two_iterations_generic(M, dd, tol2, maxrr, max_iter, k0):
    B = np.zeros((5, 5))
    B[1, 0] = B[2, 1] = B[4, 3] = 1.0
    pp, rr, xx, dl = np.eye(5)[0], np.eye(5)[3], np.zeros(5), M[3, 3]
    for j in range(2):
        w = B @ pp
        pw = pp @ M @ w
        a = dl / pw if pw != 0 else 0.0
        xx, rr = xx + a * pp, rr - a * w
        dn = max(rr @ M @ rr, 0.0)
        maxrr = max(maxrr, dn)
        if dn < dd * maxrr or dn < tol2 or dn == 0 or k0 + j == max_iter - 1:
            return "stop", j, pp, rr, xx, dl
        pp, dl = rr + (dn / dl) * pp, dn
    return "go", 2, pp, rr, xx, dl


def _two_iterations_written_out(M, dd, tol2, maxrr, max_iter, k0):
    """The device prelude's form: the coordinates' zeros written out."""
    a0 = M[3, 3] / M[0, 1] if M[0, 1] != 0 else 0.0
    dn1 = max(M[3, 3] - 2 * a0 * M[1, 3] + a0 * a0 * M[1, 1], 0.0)
    maxrr = max(maxrr, dn1)
    if dn1 < dd * maxrr or dn1 < tol2 or dn1 == 0 or k0 == max_iter - 1:
        return "stop", 0, np.eye(5)[0], np.array([0, -a0, 0, 1.0, 0]), np.array([a0, 0, 0, 0, 0.0]), M[3, 3]
    b0 = dn1 / M[3, 3]
    P, W = {0: b0, 1: -a0, 3: 1.0}, {1: b0, 2: -a0, 4: 1.0}
    pw1 = sum(P[i] * W[j] * M[i, j] for i in P for j in W)
    a1 = dn1 / pw1 if pw1 != 0 else 0.0
    R = np.array([0, -a0 - a1 * b0, a1 * a0, 1.0, -a1])
    X = np.array([a0 + a1 * b0, -a1 * a0, 0, a1, 0])
    dn2 = max(R @ M @ R, 0.0)
    maxrr = max(maxrr, dn2)
    p1 = np.array([b0, -a0, 0, 1.0, 0])
    if dn2 < dd * maxrr or dn2 < tol2 or dn2 == 0 or k0 + 1 == max_iter - 1:
        return "stop", 1, p1, R, X, dn1
    return "go", 2, R + (dn2 / dn1) * p1, R, X, dn2


def test_written_out_iterations_match_generic():
    """The device prelude writes the two coordinate iterations out with the
    known zeros of p', r', w'; on random Gram matrices (and every stop
    position) it equals the generic 5 x 5 form."""
    rng = np.random.default_rng(1)
    for t in range(500):
        Y = rng.normal(size=(5, 40))
        args = (Y @ Y.T, 0.01 * rng.random(), 0.0, 0.0, 5 if t % 3 else 2, 0 if t % 5 else 1)
        g, h = _two_iterations_generic(*args), _two_iterations_written_out(*args)
        assert g[:2] == h[:2]
        for a, b in zip(g[2:5], h[2:5]):
            assert np.allclose(a, b, rtol=1e-9, atol=1e-12)
        assert abs(g[5] - h[5]) <= 1e-9 * abs(g[5])
