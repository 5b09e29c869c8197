"""Shared input generation for the tests: seeded std::mt19937 streams drawn
by the oracle (bit-identical to the reference's own generators)."""
import numpy as np


def lattice_inputs(orc, lx, ly, beta=6.25, seed=1337):
    """(links, complex rhs, real rhs) from one seeded stream."""
    g = orc.Rng(seed)
    links = g.gauss_gauge(lx, ly, beta)
    bc = g.gaussian_complex(lx * ly)
    br = g.gaussian_real(lx * ly)
    return links, bc, br


def relerr(a, b):
    a = np.asarray(a)
    b = np.asarray(b)
    nb = np.linalg.norm(b)
    return np.linalg.norm(a - b) / (nb if nb > 0 else 1.0)
