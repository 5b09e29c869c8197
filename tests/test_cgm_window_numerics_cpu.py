"""The batched multishift solver's deferred shift updates, restated in numpy
(CPU only): the device keeps the reference's iteration (generic_cg_m.cpp:
312-600) but applies x_s -= beta_s p_s and p_s = zeta_s r + alpha_s p_s
window by window, as linear combinations of p_s at the window's start and
the window's residuals, held in a ring of kWin + 1 vectors.  This restates
the finaliser's coefficient recurrences, the window ends (every convergence
check, every kWin iterations, the stop) and the ring's slots exactly as the
kernels use them, and checks the result against the oracle's CG-M:
identical iteration count and ops_count, and per-shift solutions that differ
from the restatement's own per-iteration form (kwin = 1, the reference's
updates one by one) by at most 1e-13 -- the rounding of the window
combinations -- and from the oracle by no more than numpy's scalar
arithmetic (its complex division is not the oracle's) explains, 5e-12.

A dropped term, a wrong ring slot, a window that is not ended at a check (a
converged shift would then miss its last x updates) or a coefficient updated
in the wrong order fails it (the planted-mistake cases below).
"""
import numpy as np
import pytest

from helpers import lattice_inputs, relerr

KWIN = 10
KRING = KWIN + 1


def cgm_windowed(A, b, shifts, resid_freq, max_iter, eps, kwin=KWIN, plant=None):
    """Multishift CG with deferred shift updates; returns (X, iters, ops)."""
    n_shift = len(shifts)
    kring = kwin + 1
    ring = [None] * kring
    ring[0] = b.copy()
    p = b.copy()
    ps = [b.copy() for _ in range(n_shift)]
    xs = [np.zeros_like(b) for _ in range(n_shift)]
    vec = list(range(n_shift))
    sh = list(shifts)
    zeta = [1 + 0j] * n_shift
    zeta_p = [1 + 0j] * n_shift
    beta_s = [1 + 0j] * n_shift
    alpha_s = [0j] * n_shift
    cp = [1 + 0j] * n_shift
    ax = [0j] * n_shift
    dp = [[0j] * kwin for _ in range(n_shift)]
    ex = [[0j] * kwin for _ in range(n_shift)]
    wlen, flush = 0, False
    n_rem, n_xupd = n_shift, 0
    bsqrt = np.sqrt(np.vdot(b, b).real)
    rsq = bsqrt * bsqrt
    beta, alpha = 1 + 0j, 0j
    tol = eps * bsqrt

    def terms(kcur, wl):
        return [ring[(kcur - wl + 1 + j) % kring] for j in range(wl)]

    def apply_window(kcur, nx, np_, wl):
        rj = terms(kcur, wl)
        for q in range(nx):
            v = vec[q]
            old = ps[v].copy()
            xa = xs[v] + ax[q] * old
            for j in range(wl):
                xa = xa + ex[q][j] * rj[j]
            xs[v] = xa
            if q < np_:
                pa = cp[q] * old
                for j in range(wl):
                    pa = pa + dp[q][j] * rj[j]
                ps[v] = pa

    def beta_step(Ap):
        nonlocal beta
        bp = beta
        pap = np.vdot(p, Ap)
        beta = -rsq / pap
        for q in range(n_rem):
            z, zp = zeta[q], zeta_p[q]
            den = (beta * alpha) * (zp - z) + (zp * bp) * (1 - sh[q] * beta)
            zeta[q] = (z * zp) * bp / den
            zeta_p[q] = z
            beta_s[q] = beta * zeta[q] / zeta_p[q]

    Ap = A(p)
    beta_step(Ap)
    k = 0
    kfinal = None
    while True:
        # residual pass: the last window, then r^(k+1)
        if flush:
            apply_window(k, n_xupd, n_rem, wlen)
        r_new = ring[k % kring] + beta * Ap
        ring[(k + 1) % kring] = r_new
        rn = np.vdot(r_new, r_new).real
        # finaliser
        if flush:
            flush = False
            wlen = 0
            for q in range(n_rem):
                cp[q], ax[q] = 1 + 0j, 0j
                dp[q] = [0j] * kwin
                ex[q] = [0j] * kwin
        wl = wlen
        n_xupd = n_rem
        for q in range(n_rem):
            bq = beta_s[q]
            ax[q] = ax[q] - bq * cp[q]
            for j in range(wl):
                ex[q][j] = ex[q][j] - bq * dp[q][j]
        if k % resid_freq == 0:
            q = 0
            while q < n_rem:
                if abs(zeta[q]) * np.sqrt(rn) < tol:
                    n_rem -= 1
                    e = n_rem
                    if e != q:
                        for arr in (vec, sh, zeta, zeta_p, alpha_s, beta_s, cp, ax, dp, ex):
                            arr[e], arr[q] = arr[q], arr[e]
                        continue
                q += 1
        if n_rem == 0 or k == max_iter - 1:
            kfinal = k
            break
        al = rn / rsq
        alpha = al + 0j
        rsq = rn
        for q in range(n_rem):
            alpha_s[q] = (alpha * zeta[q]) * beta_s[q] / (zeta_p[q] * beta)
        for q in range(n_rem):
            a = alpha_s[q]
            cp[q] = a * cp[q]
            for j in range(wl):
                dp[q][j] = a * dp[q][j]
            dp[q][wl] = zeta[q] if plant != "drop_zeta" else 0j
        wlen = wl + 1
        ends = (k % resid_freq == 0) if plant != "no_check_end" else False
        if ends or wlen == kwin:
            flush = True
        k += 1
        # D^dag D on p^(k) = r^(k) + alpha p^(k-1), formed in the stencil's prologue
        slot = k % kring if plant != "ring_slot" else (k - 1) % kring
        p = ring[slot] + alpha * p
        Ap = A(p)
        beta_step(Ap)
    # the open window's x part
    rj = terms(kfinal, wlen)
    for q in range(n_xupd):
        v = vec[q]
        xa = xs[v] + ax[q] * ps[v]
        for j in range(wlen):
            xa = xa + ex[q][j] * rj[j]
        xs[v] = xa
    return xs, kfinal + 1, 1 + kfinal + n_shift


def _case(orc, L=16, mass=0.05, seed=7):
    links, bc, _ = lattice_inputs(orc, L, L, beta=6.25, seed=seed)
    A = lambda v: orc.apply("staggered_normal", v, links, mass, L, L)
    return links, bc, A


SHIFTS = 10.0 ** (-3.0 + 0.5 * np.array([0, 2, 1, 3, 5, 4, 6, 8, 7, 9, 11, 10]))


@pytest.mark.parametrize("resid_freq", [10, 3, 1, 25])
def test_windowed_cgm_matches_oracle(orc, resid_freq):
    L, mass = 16, 0.05
    links, b, A = _case(orc, L, mass)
    X, it, ops = cgm_windowed(A, b, SHIFTS, resid_freq, 4000, 1e-10)
    X1, it1, _ = cgm_windowed(A, b, SHIFTS, resid_freq, 4000, 1e-10, kwin=1)
    Xo, io, _, _ = orc.cg_m("staggered_normal", b, SHIFTS, links=links, mass=mass, lx=L, ly=L,
                            resid_freq_check=resid_freq, max_iter=4000, eps=1e-10)
    assert (it, ops) == (io.iter, io.ops_count) and it1 == it
    for k in range(SHIFTS.size):
        assert relerr(X[k], X1[k]) < 1e-13, (k, relerr(X[k], X1[k]))
        assert relerr(X[k], Xo[k]) < 5e-12, (k, relerr(X[k], Xo[k]))


@pytest.mark.parametrize("max_iter", [1, 2, 7, 23])
def test_windowed_cgm_stops_at_max_iter(orc, max_iter):
    """An unconverged stop anywhere inside a window completes every x_s."""
    L, mass = 16, 0.01
    links, b, A = _case(orc, L, mass, seed=9)
    X, it, ops = cgm_windowed(A, b, SHIFTS, 10, max_iter, 1e-10)
    X1, _, _ = cgm_windowed(A, b, SHIFTS, 10, max_iter, 1e-10, kwin=1)
    Xo, io, _, _ = orc.cg_m("staggered_normal", b, SHIFTS, links=links, mass=mass, lx=L, ly=L,
                            resid_freq_check=10, max_iter=max_iter, eps=1e-10)
    assert (it, ops) == (io.iter, io.ops_count)
    for k in range(SHIFTS.size):
        assert relerr(X[k], X1[k]) < 1e-13
        assert relerr(X[k], Xo[k]) < 5e-12


@pytest.mark.parametrize("plant", ["drop_zeta", "no_check_end", "ring_slot"])
def test_planted_mistakes_fail(orc, plant):
    """Each plausible slip in the restatement moves the solutions far beyond
    the bar above (so the bar pins the algebra)."""
    L, mass = 16, 0.05
    links, b, A = _case(orc, L, mass)
    rf = 3  # checks that do not fall on window ends
    Xo, io, _, _ = orc.cg_m("staggered_normal", b, SHIFTS, links=links, mass=mass, lx=L, ly=L,
                            resid_freq_check=rf, max_iter=4000, eps=1e-10)
    X, it, _ = cgm_windowed(A, b, SHIFTS, rf, 4000, 1e-10, plant=plant)
    worst = max(relerr(X[k], Xo[k]) for k in range(SHIFTS.size))
    assert it != io.iter or worst > 1e-6, (plant, worst)
