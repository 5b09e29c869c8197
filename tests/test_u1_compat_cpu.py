"""U(1) gauge utilities of the C++ drop-in (compat/u1_utils.h, host code in
libglx.so): a driver written against the reference interface runs the sequence
of u1_utils/u1_test.cpp:17-83 with std::mt19937(1337).

Pinned by
- the reference's own committed output of that sequence,
  u1_utils/cfg/cfg16_hot.dat (kept as tests/golden/cfg16_hot.dat): byte-equal;
- the oracle's seeded plaquettes (themselves pinned against the same file);
- the reference build, where oracle/_ref exists, for APE smearing and the
  topological charge (not restated in the oracle);
- invariants: unit plaquette 1, gauge invariance of the plaquette and of
  smearing, integer topological charge, APE with alpha = 0 is a projection.
"""
import os
import subprocess

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
GOLD = os.path.join(ROOT, "tests", "golden", "cfg16_hot.dat")
L = 16


def _betas():
    out, beta = [], 100.0
    while beta > 1e-4:  # u1_test.cpp:44 (the product sequence, not exact powers)
        out.append(beta)
        beta *= 0.1
    return out


@pytest.fixture(scope="module")
def driver_run(tmp_path_factory):
    lib = os.path.join(ROOT, "generic-linalg_amd", "libglx.so")
    if not os.path.exists(lib):
        pytest.skip("libglx.so not built")
    d = os.path.join(ROOT, "tests", "dropin")
    subprocess.run(["make", "-s", "-C", d, "u1_driver"], check=True)
    tmp = tmp_path_factory.mktemp("u1")
    hot, ape = str(tmp / "hot.dat"), str(tmp / "ape.dat")
    out = subprocess.run([os.path.join(d, "u1_driver"), hot, ape], capture_output=True, text=True, timeout=60)
    assert out.returncode == 0, out.stdout + out.stderr
    rows = {}
    gauss = []
    for line in out.stdout.splitlines():
        f = line.split()
        if f[0] == "gauss":
            gauss.append((float(f[1]), complex(float(f[2]), float(f[3]))))
        elif f[0] in ("topo",):
            rows.setdefault("topo", []).append(float(f[1]))
        elif f[0] == "ape0":
            rows.setdefault("ape0", []).append(complex(float(f[1]), float(f[2])))
        elif len(f) == 3:
            rows[f[0]] = complex(float(f[1]), float(f[2]))
        else:
            rows[f[0]] = float(f[1])
    return {"rows": rows, "gauss": gauss, "hot": hot, "ape": ape}


def test_hot_field_file_is_the_references(driver_run):
    with open(driver_run["hot"], "rb") as a, open(GOLD, "rb") as b:
        assert a.read() == b.read()


def test_plaquettes_match_oracle(driver_run, orc):
    r = driver_run["rows"]
    assert r["unit"] == 1.0 + 0j
    g = orc.Rng(1337)
    betas = _betas()
    assert [b for b, _ in driver_run["gauss"]] == betas
    for beta, p in driver_run["gauss"]:
        u = g.gauss_gauge(L, L, beta)
        assert orc.plaquette(u, L, L) == p, beta
    u = g.rand_gauge(L, L)
    assert orc.plaquette(u, L, L) == r["hot"]
    v = orc.read_gauge(GOLD, L, L)
    assert orc.plaquette(v, L, L) == r["loaded"]
    t = g.rand_trans(L, L)
    assert orc.plaquette(orc.apply_gauge_trans(v, t, L, L), L, L) == r["transformed"]
    assert abs(r["transformed"] - r["loaded"]) < 1e-14


def test_topology_and_smearing_invariants(driver_run):
    r = driver_run["rows"]
    for q in r["topo"]:
        assert abs(q - round(q)) < 1e-12
    # alpha = 0: one projection step leaves every plaquette unchanged to rounding
    for (_, p), p0 in zip(driver_run["gauss"], r["ape0"]):
        assert abs(p - p0) < 1e-14
    # smearing commutes with a gauge transform
    assert abs(r["ape3"] - r["ape3g"]) < 1e-13
    assert round(r["ape3topo"]) == round(r["ape3gtopo"])
    assert r["ape3"].real > 0.9


def test_smearing_matches_reference_build(driver_run, orc, tmp_path):
    if not orc.ref_available():
        pytest.skip("oracle/_ref not built (needs the reference sources)")
    import ctypes as C

    ref = orc.ref_lib()
    g = orc.Rng(1337, impl="ref")
    for beta in _betas():
        g.gauss_gauge(L, L, beta)
    g.rand_gauge(L, L)
    g.rand_trans(L, L)
    u = g.gauss_gauge(L, L, 4.0)
    out = np.zeros_like(u)
    ref.ref_ape_smear_u1(out.ctypes.data, u.ctypes.data, L, L, C.c_double(0.5), 3)
    path = str(tmp_path / "ref_ape.dat")
    orc.write_gauge(path, out, L, L, impl="ref")
    with open(path) as a, open(driver_run["ape"]) as b:
        assert a.read() == b.read()
    assert ref.ref_topo_u1(out.ctypes.data, L, L) == driver_run["rows"]["ape3topo"]
