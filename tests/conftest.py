"""pytest configuration: the `gpu` marker and import paths.

GPU tests call the HIP library through its C ABI and compare with the
oracle (oracle/orc.py) on the same seeded inputs.  Run them with
``pytest -m gpu`` on a machine with an MI355X; ``-m "not gpu"`` runs the CPU
suite (oracle pinning, host logic, library load/exports).
"""
import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
for p in (ROOT, os.path.join(ROOT, "oracle"), os.path.join(ROOT, "generic-linalg_amd")):
    if p not in sys.path:
        sys.path.insert(0, p)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an AMD GPU (MI355X) and the built HIP library")


@pytest.fixture(scope="session")
def glx():
    import glx as g
    g.lib()
    assert g.device_count() > 0, "no GPU visible: GPU tests need an MI355X"
    return g


@pytest.fixture(scope="session")
def orc():
    import orc as o
    o.orc_lib()
    return o


@pytest.fixture
def rows_reset(glx, monkeypatch):
    """Launch-shape table entries a test sets through glx.set_normal_rows are
    removed afterwards and the default goes back to -1 (by block count), so
    the tests after it run the library's own launch shapes.  Meanwhile the
    plain D^dag D goes to the row-marching or tiled kernel the test selects,
    not to the wave-march kernel that takes it by default."""
    glx.set_ddwave_rows(0)
    real = glx.set_normal_rows
    touched = set()

    def recording(lx, ly, esize, rows):
        if lx:
            touched.add((lx, ly, esize))
        real(lx, ly, esize, rows)

    monkeypatch.setattr(glx, "set_normal_rows", recording)
    yield touched
    for lx, ly, es in touched:
        real(lx, ly, es, -1)
    real(0, 0, 0, -1)
    glx.set_ddwave_rows(-1)
