#!/usr/bin/env python3
"""Benchmark of the MI355X solver path on its headline workload.

Workload (one GPU): CG on the 2D U(1) staggered normal operator D^dag D,
L = 1024 x 1024, m = 0.02, link angles theta ~ N(0, 0.5^2), complex Gaussian
rhs, to true relative residual 1e-10 -- plain FP64 CG against mixed-precision
CG (FP32 Krylov vectors and links, FP64 reliable updates at delta = 0.1).
Inputs are synthetic (numpy, fixed seed) and resident in HBM before timing.
The launch shape of the D^dag D kernel is autotuned for the lattice first
(outside the timed region).

A "step" is one complete mixed-precision solve from x = 0.  `value` is its
time to solution in ms (lower is better).  The JSON line also carries the
FP64 solves' times (the reference-order CG and the single-pass CG), the
speed-ups, iteration counts, the mixed solution's relative error against the
FP64 one, the roofline of the mixed solve's dominant kernel (its fused
iteration stencil) measured live with HIP events, the plain D^dag D stencil
at this size and at 4096^2 (beyond the 256 MB last-level cache), and the
reference CPU path timed on the host on a bounded sample of the same solve.

Usage: python bench.py [--gpus 1] [--steps K] [--warmup W] [--L 1024]
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "generic-linalg_amd"))

HBM_PEAK_GBS = 8000.0


def make_inputs(L: int, sigma: float, seed: int):
    rng = np.random.default_rng(seed)
    theta = rng.normal(0.0, sigma, size=2 * L * L)
    links = np.exp(1j * theta)
    b = rng.normal(size=L * L) + 1j * rng.normal(size=L * L)
    return links.astype(np.complex128), b.astype(np.complex128)


def cpu_baseline(links, b, L, mass, gpu_fp64_iters, budget_s=20.0, grad=None, batch=None, small=None):
    """Reference CPU CG (oracle/_ref, built from the reference sources; the
    oracle restatement where that build is absent), one thread, on a bounded
    sample: a fixed number of iterations of the same 1024^2 solve, scaled to
    the iteration count the FP64 solve needs.  Also, one thread each: the
    512^2 gradient (reference CG solve + D chi, the oracle's force; the force
    has no reference counterpart) and one multishift solve of the batched
    workload, projected to the whole batch by site-iterations."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import orc  # noqa: E402  (CPU baseline leg only)

    impl = "ref" if orc.ref_available() else "orc"
    # calibrate: time a few iterations, then size the sample to the budget
    n_cal = 5
    t0 = time.perf_counter()
    orc.solve("cg", "staggered_normal", b, None, links=links, mass=mass, lx=L, ly=L, max_iter=n_cal, eps=1e-10,
              impl=impl)
    t_cal = time.perf_counter() - t0
    per_it = t_cal / (n_cal + 3)  # n iterations + 3 extra operator applications
    n_it = max(5, min(gpu_fp64_iters, int(budget_s / per_it) - 3))
    t0 = time.perf_counter()
    _, info = orc.solve("cg", "staggered_normal", b, None, links=links, mass=mass, lx=L, ly=L, max_iter=n_it,
                        eps=1e-10, impl=impl)
    dt = time.perf_counter() - t0
    per_it = dt / (info.iter + 3)
    projected_ms = per_it * (gpu_fp64_iters + 3) * 1e3
    out = {
        "value": round(projected_ms, 1),
        "unit": "ms (FP64 CG time to 1e-10, projected)",
        "cores": 1,
        "kind": "reference" if impl == "ref" else "port",
        "sample": f"{info.iter} CG iterations of the same {L}x{L} solve on one host core "
                  f"({dt:.2f} s, {per_it * 1e3:.1f} ms/iteration), scaled to the {gpu_fp64_iters} iterations "
                  f"the FP64 solve takes",
        "cpu_model": cpu_model(),
    }
    if grad is not None:
        Lg, mg = grad["L"], grad["mass"]
        rng = np.random.default_rng(21)
        theta = rng.normal(0.0, 0.4, 2 * Lg * Lg)
        phi = rng.normal(size=Lg * Lg) + 1j * rng.normal(size=Lg * Lg)
        lk = np.exp(1j * theta)
        t0 = time.perf_counter()
        chi, gi = orc.solve("cg", "staggered_normal", phi, None, links=lk, mass=mg, lx=Lg, ly=Lg, max_iter=100000,
                            eps=1e-10, impl=impl)
        psi = orc.apply("staggered", chi, lk, mg, Lg, Lg, impl=impl)
        orc.force(chi, psi, lk, mg, Lg, Lg)
        out["gradient_ms"] = round((time.perf_counter() - t0) * 1e3, 1)
        out["gradient_sample"] = (f"full {Lg}x{Lg} gradient on one core: {impl} CG ({gi.iter} iterations) + D chi, "
                                  f"the oracle's force")
    if small is not None:
        gpu, inputs = small
        L1, b1 = inputs["C1"]
        L2, lk2, b2 = inputs["C2"]
        for name, solver, op, mass, bb, lk, Ls in (("C1_laplace_cg", "cg", "laplace_real", 0.01, b1, None, L1),
                                                   ("C2_normal_cg", "cg", "staggered_normal", 0.05, b2, lk2, L2),
                                                   ("C2_bicgstab_D", "bicgstab", "staggered", 0.05, b2, lk2, L2)):
            t0 = time.perf_counter()
            xc, ci = orc.solve(solver, op, bb, None, links=lk, mass=mass, lx=Ls, ly=Ls, max_iter=100000,
                               eps=1e-10, impl=impl)
            dtc = time.perf_counter() - t0
            xg = gpu[name]["x"]
            out[name] = {"ms": round(dtc * 1e3, 2), "iters": ci.iter,
                         "x_relerr_gpu_vs_cpu": float(np.linalg.norm(xg - xc) / np.linalg.norm(xc))}
    if batch is not None:
        Lb, mb = 128, 0.05
        rng = np.random.default_rng(8)
        lk = np.exp(1j * rng.normal(0.0, 0.4, 2 * Lb * Lb))
        bb = rng.normal(size=Lb * Lb) + 1j * rng.normal(size=Lb * Lb)
        shifts = 10.0 ** (-3.0 + 0.5 * np.arange(12))
        t0 = time.perf_counter()
        _, ci, _, _ = orc.cg_m("staggered_normal", bb, shifts, links=lk, mass=mb, lx=Lb, ly=Lb,
                               resid_freq_check=10, max_iter=100000, eps=1e-10, impl=impl)
        dtb = time.perf_counter() - t0
        per_site_it = dtb / (Lb * Lb * ci.iter)
        out["batch_s"] = round(per_site_it * batch["site_iters"], 1)
        out["batch_sample"] = (f"one {Lb}x{Lb} 12-shift CG-M on one core ({impl}, {ci.iter} iterations, {dtb:.2f} s), "
                               f"projected to the batch's {batch['site_iters']:.3g} site-iterations")
    return out


def cpu_model():
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip() + f" ({os.cpu_count()} logical CPUs visible)"
    except OSError:
        pass
    return None


def stencil_beyond_mall(glx, L=4096):
    """FP32 D^dag D (FP32 links) at 4096^2: 128 MiB per vector and 256 MiB of
    links, beyond the 256 MB last-level cache, so the figure is HBM-bound.
    The launch shape of the operator alone is tuned here over rows per block
    (the autotuner tunes the CG iteration kernel, which stages more data)."""
    V = L * L
    rng = np.random.default_rng(5)
    links = np.exp(1j * rng.normal(0.0, 0.5, size=2 * V)).astype(np.complex128)
    lat = glx.Lattice(links, L, L, f32=True, f16=False)
    x = (rng.normal(size=V) + 1j * rng.normal(size=V)).astype(np.complex64)
    d_in = glx.DeviceArray(x)
    d_out = glx.DeviceArray(np.zeros_like(x))
    best = None
    for rows in (32, 64, 96, 128):
        glx.set_normal_rows(L, L, 8, rows)
        ms = glx.time_apply(lat, "staggered_normal", 0.02, d_out, d_in, reps=50, fp32=True,
                            link_prec=glx.LINKS_F32)
        if best is None or ms < best[1]:
            best = (rows, ms)
    rows, ms = best
    glx.set_normal_rows(L, L, 8, rows)
    gbs = 32 * V / (ms * 1e-3) / 1e9
    del d_in, d_out
    # the same operator in FP64 (64 B per site: 1 GiB per launch)
    x64 = x.astype(np.complex128)
    d_in, d_out = glx.DeviceArray(x64), glx.DeviceArray(np.zeros_like(x64))
    best64 = None
    for r64 in (32, 64, 128):
        glx.set_normal_rows(L, L, 16, r64)
        m64 = glx.time_apply(lat, "staggered_normal", 0.02, d_out, d_in, reps=30)
        if best64 is None or m64 < best64[1]:
            best64 = (r64, m64)
    glx.set_normal_rows(L, L, 16, best64[0])
    gbs64 = 64 * V / (best64[1] * 1e-3) / 1e9
    del lat, d_in, d_out
    glx.release_cache()
    return {"L": L, "rows": rows, "fp32_ms": ms, "fp32_GBs": round(gbs, 1),
            "frac": round(gbs / HBM_PEAK_GBS, 4), "fp64_rows": best64[0], "fp64_ms": best64[1],
            "fp64_GBs": round(gbs64, 1), "fp64_frac": round(gbs64 / HBM_PEAK_GBS, 4)}


def small_configs(glx, reps=5):
    """The reference's small workloads on the GPU (FP64, reference order of
    operations): C1 real Laplace 64^2 (diagonal 4 + 0.01) with CG; C2
    staggered 256^2 (m = 0.05, theta ~ N(0, 0.4^2)) with CG on D^dag D and
    BiCGStab on D.  Host-timed solves on device-resident vectors (best of
    `reps`); C1 is latency-bound, so no roofline fraction is given for it."""
    out = {}
    rng = np.random.default_rng(31)
    L1 = 64
    lat1 = glx.Lattice(np.ones(2 * L1 * L1, np.complex128), L1, L1, f32=False, f16=False)
    b1 = rng.standard_normal(L1 * L1)
    rng2 = np.random.default_rng(32)
    L2 = 256
    links2 = np.exp(1j * rng2.normal(0.0, 0.4, 2 * L2 * L2))
    b2 = rng2.standard_normal(L2 * L2) + 1j * rng2.standard_normal(L2 * L2)
    lat2 = glx.Lattice(links2, L2, L2, f32=False, f16=False)
    cases = (("C1_laplace_cg", "cg", lat1, "laplace_real", 0.01, b1),
             ("C2_normal_cg", "cg", lat2, "staggered_normal", 0.05, b2),
             ("C2_bicgstab_D", "bicgstab", lat2, "staggered", 0.05, b2))
    for name, solver, lat, op, mass, b in cases:
        bd = glx.DeviceArray(b)
        x = glx.DeviceArray(np.zeros_like(b))
        best, info = 1e9, None
        for _ in range(reps):
            x.upload(np.zeros_like(b))
            glx.synchronize()
            t0 = time.perf_counter()
            info = glx.dev_solve(solver, lat, op, mass, x, bd, max_iter=100000, eps=1e-10)
            glx.synchronize()
            best = min(best, time.perf_counter() - t0)
        out[name] = {"ms": round(best * 1e3, 3), "iters": info.iter, "success": info.success,
                     "L": lat.lx, "mass": mass, "x": x.download()}
    return out, {"C1": (L1, b1), "C2": (L2, links2, b2)}


def cstencil_traffic(L: int, nc: int):
    """HBM bytes per launch of k_cstencil for one bench shape, from the
    committed rocprofv3 FETCH_SIZE / WRITE_SIZE summary; None if absent."""
    path = os.path.join(ROOT, "profiles", "k_cstencil_pmc.txt")
    if not os.path.exists(path):
        return None
    hit = False
    with open(path) as f:
        for line in f:
            if line.startswith(f"L={L} nc={nc} "):
                hit = True
            elif hit and "traffic_bytes_per_launch:" in line:
                return float(line.split(":", 1)[1])
    return None


def general_stencil(glx):
    """General stencil_2d (coarse_stencil.cpp:12-392) on the device: the
    per-launch bytes are the coefficients (n_terms nc^2 16 B per site, read
    once) plus the input and output vectors (2 nc 16 B per site).  Two shapes
    of multigrid coarse operators: nc = 8 with two-link terms on 128^2
    (218 MB of coefficients, beyond the last-level cache) and nc = 16
    one-link on 64^2."""
    out = []
    rng = np.random.default_rng(9)
    for L, nc, two in ((128, 8, True), (64, 16, False)):
        V = L * L
        blk = V * nc * nc
        g = lambda k: (rng.standard_normal(k * blk) + 1j * rng.standard_normal(k * blk))
        st = glx.StencilNc(L, L, nc, g(1), g(4), g(8) if two else None, 0.1)
        h = glx.CStencilDev(st)
        del st
        v = (rng.standard_normal(V * nc) + 1j * rng.standard_normal(V * nc))
        d_in, d_out = glx.DeviceArray(v), glx.DeviceArray(np.zeros_like(v))
        ms = h.time(d_out, d_in, "all", reps=50)
        nterms = 13 if two else 5
        by = V * (nterms * nc * nc * 16 + 2 * nc * 16)
        gbs = by / (ms * 1e-3) / 1e9
        out.append({"L": L, "nc": nc, "two_link": two, "ms": ms, "bytes": by, "GBs": round(gbs, 1),
                    "frac": round(gbs / HBM_PEAK_GBS, 4), "traffic": cstencil_traffic(L, nc)})
        h.close()
        del d_in, d_out
    glx.release_cache()
    return out


def gauge_generation(glx, L=512, reps=10):
    """Gaussian U(1) field generated in HBM (Philox4x32-10, Box-Muller) with
    its FP32/FP16 copies, and the device mean plaquette, for 512^2."""
    lat = glx.Lattice.gaussian(L, L, 6.25, seed=1)
    glx.synchronize()
    t0 = time.perf_counter()
    for r in range(reps):
        lat.regenerate(6.25, seed=r + 2)
    glx.synchronize()
    gen_ms = (time.perf_counter() - t0) * 1e3 / reps
    t0 = time.perf_counter()
    for r in range(reps):
        p = lat.plaquette()
    plaq_ms = (time.perf_counter() - t0) * 1e3 / reps
    lat.close()
    return {"L": L, "beta": 6.25, "generate_ms": round(gen_ms, 4), "plaquette_ms": round(plaq_ms, 4),
            "plaquette": [p.real, p.imag], "note": "host-timed, each call synchronises"}


def gradient_timing(glx, L=512, mass=0.05, reps=3):
    """Gradient of S = phi^dag (D^dag D)^-1 phi on the 512^2 workload
    (theta ~ N(0, 0.4^2), phi complex Gaussian): adjoint solve to 1e-10 +
    D chi + per-link force, FP64 and FP32-bulk, gradient left in HBM."""
    rng = np.random.default_rng(21)
    theta = rng.normal(0.0, 0.4, 2 * L * L)
    phi = rng.normal(size=L * L) + 1j * rng.normal(size=L * L)
    lat = glx.Lattice(np.exp(1j * theta), L, L)
    glx.autotune_normal(lat, reps=20)
    dphi = glx.DeviceArray(phi)
    g = glx.DeviceArray(nbytes=2 * L * L * 8, dtype=np.float64, shape=(2 * L * L,))
    out = {"L": L, "mass": mass, "tol": 1e-10}
    ref = None
    for name, prec in (("fp64", 0), ("fp32_bulk", glx.LINKS_F32)):
        glx.staggered_grad(lat, mass, dphi, precision=prec, grad=g, download=False)  # warm-up
        ts = []
        for _ in range(reps):
            glx.synchronize()
            t0 = time.perf_counter()
            r = glx.staggered_grad(lat, mass, dphi, precision=prec, grad=g, download=False)
            glx.synchronize()
            ts.append(time.perf_counter() - t0)
        gv = g.download()
        if ref is None:
            ref = gv
        out[name] = {"ms": round(float(np.median(ts)) * 1e3, 3), "iters": r.info.iter, "dS_dm": r.dS_dm,
                     "relerr_vs_fp64": float(np.linalg.norm(gv - ref) / np.linalg.norm(ref))}
    del lat
    glx.release_cache()
    return out


def batch_multishift(glx, n_lat=48, rhs=4, seq=True):
    """48 staggered lattices, L uniform on {64, ..., 512}, m ~ U[0.01, 0.1],
    theta ~ N(0, 0.4^2), 4 complex Gaussian right-hand sides each, 12 shifts
    sigma_k = 10^(-3 + 0.5 k), tol 1e-10, resid_freq_check 10: the 192
    multishift solves in one batched call (launch shapes autotuned per L),
    and the same solves one call each for comparison."""
    rng = np.random.default_rng(4)
    Ls = rng.choice([64, 128, 192, 256, 384, 512], size=n_lat)
    masses = rng.uniform(0.01, 0.1, size=n_lat)
    shifts = 10.0 ** (-3.0 + 0.5 * np.arange(12))
    tuned, systems = {}, []
    for i, L in enumerate(Ls):
        L = int(L)
        lat = glx.Lattice(np.exp(1j * rng.normal(0.0, 0.4, 2 * L * L)), L, L, f32=False, f16=False)
        if L not in tuned:
            tuned[L] = glx.autotune_normal(lat, reps=10)["rows64"]
        for _ in range(rhs):
            b = rng.normal(size=L * L) + 1j * rng.normal(size=L * L)
            xs = [glx.DeviceArray(nbytes=L * L * 16, shape=(L * L,)) for _ in shifts]
            systems.append((lat, float(masses[i]), glx.DeviceArray(b), xs))
    glx.dev_cg_m_batched(systems[:8], shifts, max_iter=20)  # warm-up (code paths, allocations)
    glx.synchronize()
    t0 = time.perf_counter()
    infos = glx.dev_cg_m_batched(systems, shifts, resid_freq_check=10, max_iter=100000, eps=1e-10)
    glx.synchronize()
    t_batch = time.perf_counter() - t0
    site_iters = float(sum(lat.lx * lat.ly * inf.iter for (lat, _, _, _), inf in zip(systems, infos)))
    out = {"systems": len(systems), "lattices": n_lat, "rhs": rhs, "shifts": 12,
           "sites_total": int(sum(int(L) * int(L) for L in Ls) * rhs),
           "batch_s": round(t_batch, 3), "all_success": all(i.success for i in infos),
           "iters_max": max(i.iter for i in infos), "iters_sum": sum(i.iter for i in infos),
           "site_iters": site_iters,
           "rows_by_L": {str(k): v for k, v in sorted(tuned.items())}}
    if seq:
        t0 = time.perf_counter()
        for lat, m, db, xs in systems:
            glx.dev_cg_m(lat, "staggered_normal", m, xs, db, shifts, resid_freq_check=10, max_iter=100000,
                         eps=1e-10)
        glx.synchronize()
        out["one_call_each_s"] = round(time.perf_counter() - t0, 3)
    del systems
    glx.release_cache()
    return out


def eo_solves(glx, lat, b, mass, reps=3):
    """D x = b on the same lattice by the even/odd pipeline on compacted
    vectors (prepare, CG on m^2 - D_eo D_oe to 1e-10, reconstruct): FP64
    single-pass and mixed precision, with the full residual |D x - b| / |b|."""
    out = {}
    db = glx.DeviceArray(b)
    dx = glx.DeviceArray(np.zeros_like(b))
    bn = float(np.linalg.norm(b))
    for method in ("cg", "cg_fused", "mixed"):
        glx.dev_solve_staggered_eo(lat, mass, dx, db, method, eps=1e-10)  # warm-up
        ts = []
        for _ in range(reps):
            glx.synchronize()
            t0 = time.perf_counter()
            info, r2 = glx.dev_solve_staggered_eo(lat, mass, dx, db, method, eps=1e-10)
            glx.synchronize()
            ts.append(time.perf_counter() - t0)
        out[method] = {"ms": round(float(np.median(ts)) * 1e3, 3), "iters": info.iter,
                       "full_relres": float(np.sqrt(r2) / bn)}
    return out


def stencil_timing(glx, lat, b, mass, L, reps=200):
    """The operators alone at L^2 and 2 L^2: D (the tiled kernel) and the fused
    D^dag D, FP64 and FP32 (and FP16 links), HIP events around `reps`
    back-to-back launches.  A lone D^dag D application takes its own best
    rows per block (the table's entry is tuned for the CG iteration kernel,
    whose prologue and epilogue change the balance); the table is restored
    afterwards.  Algorithmic bytes per application: V (2 s_v + 2 s_u [+ 4])."""
    out = {"op": "D and fused D^dag D, algorithmic bytes V*(2 s_v + 2 s_u [+ s_scale])"}

    def measure(lt, vb, n, key):
        V = n * n
        d_in, d_out = glx.DeviceArray(vb), glx.DeviceArray(np.zeros_like(vb))
        v32 = vb.astype(np.complex64)
        i32, o32 = glx.DeviceArray(v32), glx.DeviceArray(np.zeros_like(v32))
        res = {"L": n}
        for op, tag in (("staggered", "D"), ("staggered_normal", "DdD")):
            saved = (glx.get_normal_rows(n, n, 16), glx.get_normal_rows(n, n, 8))
            cand = (0, 8, 16, 32, 64) if tag == "DdD" else (None,)
            best = {}
            for rows in cand:
                if rows is not None:
                    glx.set_normal_rows(n, n, 16, rows)
                    glx.set_normal_rows(n, n, 8, rows)
                for prec, fp32, lp, by in (("fp64", False, glx.LINKS_F32, 64), ("fp32", True, glx.LINKS_F32, 32),
                                          ("fp16link", True, glx.LINKS_F16, 28)):
                    di, do = (i32, o32) if fp32 else (d_in, d_out)
                    ms = glx.time_apply(lt, op, mass, do, di, reps=reps, fp32=fp32, link_prec=lp)
                    if prec not in best or ms < best[prec][0]:
                        best[prec] = (ms, rows)
            glx.set_normal_rows(n, n, 16, saved[0])
            glx.set_normal_rows(n, n, 8, saved[1])
            for prec, (ms, rows) in best.items():
                by = {"fp64": 64, "fp32": 32, "fp16link": 28}[prec] * V
                gbs = by / (ms * 1e-3) / 1e9
                res[f"{tag}_{prec}"] = {"us": round(ms * 1e3, 2), "GBs": round(gbs, 1),
                                        "frac": round(gbs / HBM_PEAK_GBS, 4), "rows": rows}
        out[key] = res

    measure(lat, b, L, "at_L")
    L2 = 2 * L
    links2, b2 = make_inputs(L2, 0.5, seed=1338)
    lat2 = glx.Lattice(links2, L2, L2, f32=True, f16=True)
    measure(lat2, b2, L2, "at_2L")
    del lat2
    return out


def traffic_from_profile(kernel: str, links: str):
    """HBM bytes per launch of the dominant kernel as measured by rocprofv3
    FETCH_SIZE / WRITE_SIZE (scaled by a copy kernel of known size), read from
    the committed summary for this workload; None if there is none."""
    path = os.path.join(ROOT, "profiles", f"{kernel}_fp32_{links}_L1024_pmc.txt")
    if not os.path.exists(path):
        return None
    vals = {}
    with open(path) as f:
        for line in f:
            if line.startswith("traffic_bytes_per_launch:"):
                vals["t"] = float(line.split(":", 1)[1])
    return vals.get("t")


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--L", type=int, default=1024)
    ap.add_argument("--mass", type=float, default=0.02)
    ap.add_argument("--sigma", type=float, default=0.5)
    ap.add_argument("--links", choices=["f32", "f16"], default="f32")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-big", action="store_true", help="skip the 4096^2 stencil measurement")
    ap.add_argument("--no-grad", action="store_true", help="skip the 512^2 gradient timing")
    ap.add_argument("--no-eo", action="store_true", help="skip the even/odd D x = b solves")
    ap.add_argument("--no-batch", action="store_true", help="skip the batched multishift workload")
    ap.add_argument("--no-cstencil", action="store_true", help="skip the general stencil_2d timing")
    ap.add_argument("--sstep-rows", type=int, default=0,
                    help="rows per block of the mixed solver's s-step kernel (0: autotune; a profiler changes the "
                         "timings the autotuner compares, so profiled runs pin the shape of the unprofiled one)")
    ap.add_argument("--copy-mib", type=int, default=0,
                    help="first run 3 device-to-device copies of this many MiB (a plain copy kernel of known size, "
                         "for scaling rocprofv3 byte counters)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(json.dumps({"error": "this build measures one GPU: the solver path has no multi-GPU code; "
                                   "run with --gpus 1"}))
        return 2

    import glx

    if args.copy_mib > 0:
        n = args.copy_mib << 16  # complex128 elements (16 bytes)
        src = glx.DeviceArray(np.ones(n, np.complex128))
        dst = glx.DeviceArray(np.zeros(n, np.complex128))
        for _ in range(3):
            glx.copy_device(dst, src)
        glx.synchronize()
        del src, dst

    L, mass = args.L, args.mass
    V = L * L
    links, b = make_inputs(L, args.sigma, seed=1337)
    lat = glx.Lattice(links, L, L, f32=True, f16=True)
    glx.set_tuning(group=16, use_graph=False)
    tuned = glx.autotune_normal(lat, reps=40)
    if args.sstep_rows > 0:
        glx.set_sstep_rows(L, L, args.sstep_rows)
        tuned_ss = {"rows": args.sstep_rows}
    else:
        tuned_ss = glx.autotune_sstep(lat, reps=40)
    d_b = glx.DeviceArray(b)
    d_x = glx.DeviceArray(np.zeros_like(b))
    d_x64 = glx.DeviceArray(np.zeros_like(b))
    zero = np.zeros_like(b)
    link_prec = glx.LINKS_F16 if args.links == "f16" else glx.LINKS_F32

    def solve_fp64():
        d_x64.upload(zero)
        glx.synchronize()
        t0 = time.perf_counter()
        info = glx.dev_solve("cg", lat, "staggered_normal", mass, d_x64, d_b, max_iter=100000, eps=1e-10)
        glx.synchronize()
        return time.perf_counter() - t0, info

    def solve_fused():
        d_x64.upload(zero)
        glx.synchronize()
        t0 = time.perf_counter()
        info = glx.dev_cg_fused(lat, "staggered_normal", mass, d_x64, d_b, max_iter=100000, eps=1e-10)
        glx.synchronize()
        return time.perf_counter() - t0, info

    def solve_mixed(sstep=2):
        glx.set_sstep(sstep)
        d_x.upload(zero)
        glx.synchronize()
        t0 = time.perf_counter()
        info = glx.dev_cg_mixed(lat, "staggered_normal", mass, d_x, d_b, max_iter=100000, eps=1e-10, delta=0.1,
                                link_prec=link_prec)
        glx.synchronize()
        glx.set_sstep()
        return time.perf_counter() - t0, info

    for _ in range(args.warmup):
        solve_fp64()
        solve_fused()
        solve_mixed(1)
        solve_mixed()
    tfu = []
    for _ in range(max(1, args.steps)):
        dt, infofu = solve_fused()
        tfu.append(dt)
    xfu = d_x64.download()
    t64 = []
    for _ in range(max(1, args.steps)):
        dt, info64 = solve_fp64()
        t64.append(dt)
    tm1 = []
    for _ in range(max(1, args.steps)):
        dt, infom1 = solve_mixed(1)
        tm1.append(dt)
    tmx = []
    for _ in range(args.steps):
        dt, infomx = solve_mixed()
        tmx.append(dt)
    ms64 = float(np.median(t64)) * 1e3
    msfu = float(np.median(tfu)) * 1e3
    msmx = float(np.median(tmx)) * 1e3
    msm1 = float(np.median(tm1)) * 1e3
    best64 = min(ms64, msfu)
    x64 = d_x64.download()
    xmx = d_x.download()
    err_vs_fp64 = float(np.linalg.norm(xmx - x64) / np.linalg.norm(x64))

    # live roofline of the mixed solve's dominant kernel: the s-step launch
    # (two CG iterations: p, r, x from the previous basis, then A p, A^2 p,
    # A r and twelve FP64 partial sums), HIP events around back-to-back
    # launches on its own stream; the one-iteration kernel beside it
    it_ms, it_bps = glx.time_cg_kernel(lat, mass, prec=2 + link_prec, reps=200)
    it1_ms, it1_bps = glx.time_cg_kernel(lat, mass, prec=link_prec, reps=200)
    it64_ms, it64_bps = glx.time_cg_kernel(lat, mass, prec=0, reps=200)
    it_gbs = it_bps * V / (it_ms * 1e-3) / 1e9
    it64_gbs = it64_bps * V / (it64_ms * 1e-3) / 1e9

    stencil = stencil_timing(glx, lat, b, mass, L)
    big = None
    if not args.no_big:
        big = stencil_beyond_mall(glx)

    eo = None if args.no_eo else eo_solves(glx, lat, b, mass)
    grad = None if args.no_grad else gradient_timing(glx)
    batch = None if args.no_batch else batch_multishift(glx)
    cst = None if args.no_cstencil else general_stencil(glx)
    small = small_configs(glx)
    gauge = None if args.no_cstencil else gauge_generation(glx)

    cpu = None
    if not args.no_cpu:
        try:
            cpu = cpu_baseline(links, b, L, mass, info64.iter, grad=grad, batch=batch, small=small)
        except Exception as e:  # the baseline is reported, never required
            cpu = {"value": None, "unit": "ms", "cores": 1, "kind": "port", "sample": f"failed: {e}"}

    out = {
        "metric": "CG time-to-1e-10 & stencil HBM GB/s, 2D U(1) staggered: FP64 vs mixed",
        "value": round(msmx, 3),
        "unit": "ms",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(msmx, 3),
        "higher_is_better": False,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": f"fp32 bulk ({args.links} links) + fp64 reliable updates",
        "data": "synthetic (numpy seed 1337): theta ~ N(0, 0.5^2), complex Gaussian rhs",
        "config": {
            "workload": f"CG on staggered D^dag D, L={L}x{L}, m={mass}, theta~N(0,{args.sigma}^2), "
                        f"true relres 1e-10; FP64 CG vs mixed (delta 0.1)",
            "L": L, "mass": mass, "tol": 1e-10, "delta": 0.1,
        },
        "fp64_ms": round(ms64, 3),
        "fp64_fused_ms": round(msfu, 3),
        "speedup_vs_fp64": round(ms64 / msmx, 3),
        "speedup_basis": "plain FP64 GPU CG in the reference's order of operations (glx_dev_solve)",
        "speedup_vs_fp64_fused": round(msfu / msmx, 3),
        "mixed_form": (f"s-step, two CG iterations per launch, {tuned_ss['rows']} rows per block "
                       f"({'set' if args.sstep_rows > 0 else 'autotuned'})"
                       if tuned_ss["rows"] > 0 else "one CG iteration per launch (autotuned)"),
        "mixed_single_pass_ms": round(msm1, 3),
        "mixed_single_pass_iters": infom1.iter,
        "fp64_iters": info64.iter,
        "fp64_fused_iters": infofu.iter,
        "fp64_fused_relerr_vs_fp64": float(np.linalg.norm(xfu - x64) / np.linalg.norm(x64)),
        "mixed_iters": infomx.iter,
        "mixed_reliable_updates": infomx.reliable_updates,
        "mixed_true_relres": float(np.sqrt(infomx.resSq) / np.linalg.norm(b)),
        "fp64_true_relres": float(np.sqrt(info64.resSq) / np.linalg.norm(b)),
        "mixed_relerr_vs_fp64": err_vs_fp64,
        "normal_rows": {"fp64": tuned["rows64"], "fp32": tuned["rows32"]},
        "stencil": dict(stencil, beyond_mall=big),
        "cg_kernel_fp64": {"ms_per_launch": it64_ms, "bytes_per_site": it64_bps, "GBs": round(it64_gbs, 1)},
        "cg_kernel_fp32_single_pass": {"ms_per_launch": it1_ms, "bytes_per_site": it1_bps,
                                       "GBs": round(it1_bps * V / (it1_ms * 1e-3) / 1e9, 1)},
        "roofline": {
            "bound": "hbm",
            "kernel": f"k_sscg<2, Links{args.links.upper()}> (s-step mixed CG launch = two CG iterations: "
                      f"p, r, x from the previous basis + A p, A^2 p, A r + 12 partial sums)",
            "achieved": round(it_gbs, 1),
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": round(it_gbs / HBM_PEAK_GBS, 4),
            "traffic": traffic_from_profile("k_sscg", args.links),
            "bytes_per_launch": it_bps * V,
            "bytes_per_site": it_bps,
            "ms_per_launch": it_ms,
        },
        "eo_solve_D": eo,
        "gradient": grad,
        "batch_multishift": batch,
        "general_stencil": cst,
        "small_configs": {k: {kk: vv for kk, vv in v.items() if kk != "x"} for k, v in small[0].items()},
        "gauge_generation": gauge,
        "cpu_baseline": cpu,
    }
    print(json.dumps(out))
    return 0


if __name__ == "__main__":
    sys.exit(main())
