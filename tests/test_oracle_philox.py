"""The oracle's restatement of the library's device gauge-field generator
(Philox4x32-10 + Box-Muller; no reference counterpart, the reference draws
std::mt19937 on the host).  Pinned by the published Random123 known-answer
vectors for philox4x32-10 and by the distribution the fields must have:
theta ~ N(0, 1/beta), uniform phases at beta = 0, and the closed-form mean
plaquette exp(-2/beta) of independent Gaussian link angles."""
import numpy as np
import pytest
from scipy import stats

# Random123 kat_vectors, philox4x32 with 10 rounds: (counter, key) -> output
KAT = [
    ([0, 0, 0, 0], [0, 0], [0x6627E8D5, 0xE169C58D, 0xBC57AC4C, 0x9B00DBD8]),
    ([0xFFFFFFFF] * 4, [0xFFFFFFFF] * 2, [0x408F276D, 0x41C83B0E, 0xA20BC7C6, 0x6D5451FD]),
    ([0x243F6A88, 0x85A308D3, 0x13198A2E, 0x03707344], [0xA4093822, 0x299F31D0],
     [0xD16CFE09, 0x94FDCCEB, 0x5001E420, 0x24126EA1]),
]


@pytest.mark.parametrize("ctr,key,want", KAT)
def test_philox_known_answers(orc, ctr, key, want):
    assert orc.philox4x32_10(ctr, key) == want


def test_gaussian_phases(orc):
    L, beta = 256, 4.0
    th = np.angle(orc.philox_gauge(L, L, beta, seed=0x1234_5678_9ABC, stream=3))
    n = th.size
    assert abs(th.mean()) < 5 * 0.5 / np.sqrt(n)
    assert abs(th.var() - 1 / beta) < 5 * (1 / beta) * np.sqrt(2 / n)
    assert stats.kstest(th * np.sqrt(beta), "norm").pvalue > 1e-4
    # x and y links of a site come from one Box-Muller pair: uncorrelated
    assert abs(np.corrcoef(th[0::2], th[1::2])[0, 1]) < 5 / np.sqrt(n / 2)
    # |beta| is used, as by gauss_gauge_u1 (u1_utils.cpp:95)
    assert np.array_equal(orc.philox_gauge(8, 6, -beta, 7), orc.philox_gauge(8, 6, beta, 7))


def test_hot_phases(orc):
    th = np.angle(orc.philox_gauge(256, 256, 0.0, seed=99))
    assert stats.kstest((th + np.pi) / (2 * np.pi), "uniform").pvalue > 1e-4


def test_mean_plaquette_closed_form(orc):
    # four independent N(0, 1/beta) angles: <cos(sum)> = exp(-2/beta), <sin> = 0
    L = 256
    for beta in (10.0, 2.0):
        p = orc.plaquette(orc.philox_gauge(L, L, beta, seed=5), L, L)
        sd = np.sqrt((1 - np.exp(-4 / beta)) / 2) + 0.05
        assert abs(p.real - np.exp(-2 / beta)) < 5 * sd / L
        assert abs(p.imag) < 5 * sd / L


def test_counter_addressing(orc):
    # a field depends on (seed, stream) and the site index only
    a = orc.philox_gauge(16, 4, 3.0, seed=11, stream=2)
    b = orc.philox_gauge(8, 8, 3.0, seed=11, stream=2)
    assert np.array_equal(a, b)
    assert not np.array_equal(a, orc.philox_gauge(16, 4, 3.0, seed=11, stream=3))
    assert not np.array_equal(a, orc.philox_gauge(16, 4, 3.0, seed=12, stream=2))
    assert np.allclose(np.abs(a), 1.0, atol=1e-15)
