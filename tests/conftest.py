import os
import subprocess
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PKG = os.path.join(ROOT, "ab-initio-md_amd")
for p in (ROOT, PKG):
    if p not in sys.path:
        sys.path.insert(0, p)

BOHR = 0.529177210903
ANG = 1.0 / BOHR

# geometries in angstrom (the reference's test molecules, tests/qc_reference.py:35-74,
# and its example files)
WATER = (["O", "H", "H"], np.array([[0.0, 0.0, 0.1173], [0.0, 0.7572, -0.4692], [0.0, -0.7572, -0.4692]]))
WATER_STO3G_MIN = (["O", "H", "H"], np.array([[0.0, 0.0, 0.071151], [0.0, 0.758081, -0.564652],
                                              [0.0, -0.758081, -0.564652]]))
METHYL = (["C", "H", "H", "H"], np.array([[0.0, 0.0, 0.0], [1.072562, -0.000009, -0.003312],
                                          [-0.536289, 0.928862, -0.003310], [-0.536274, -0.928853, 0.006622]]))
METHANOL = (["C", "O", "H", "H", "H", "H"], np.array([
    [-0.030126, 0.716892, 0.0], [-0.022276, -0.682649, 0.0], [-1.064781, 1.030232, 0.0],
    [0.452769, 1.127616, 0.883547], [0.452769, 1.127616, -0.883547], [0.871783, -0.992697, 0.0]]))
HCL = (["H", "Cl"], np.array([[0.0, 0.0, -1.2027], [0.0, 0.0, 0.0723]]))
AMMONIA = (["N", "H", "H", "H"], np.array([[0.0, 0.0, 0.1162], [0.0, 0.9377, -0.2711],
                                           [0.8121, -0.4689, -0.2711], [-0.8121, -0.4689, -0.2711]]))


def bohr(mol, distort=0.0, seed=0):
    sym, x = mol
    pos = np.array(x) * ANG
    if distort:
        pos = pos + distort * np.random.default_rng(seed).uniform(-1.0, 1.0, pos.shape)
    return list(sym), pos


def water_dimer():
    s, x = bohr(WATER)
    return s * 2, np.concatenate([x, x + np.array([2.8, 0.3, 0.1]) * ANG])


def cluster(name):
    from aimd_amd.md import read_xyz
    sym, x = read_xyz(os.path.join(ROOT, "tests", "golden", name))
    return sym, x * ANG


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs a ROCm GPU (MI355X) and the built engine")


def _have(path):
    return os.path.exists(path)


@pytest.fixture(scope="session")
def oracle_lib():
    from oracle import oracle as O
    O.build()
    O.set_threads(min(8, os.cpu_count() or 1))
    return O


def _harness_lib(name):
    lib = os.path.join(ROOT, "tests", "harness", name)
    if not _have(lib):
        subprocess.run(["make", "-C", os.path.join(ROOT, "tests", "harness"), "-j8", name], check=True)
    from tests.harness import hostcheck
    return hostcheck.Harness(lib)


@pytest.fixture(scope="session")
def harness():
    return _harness_lib("libhostcheck.so")


@pytest.fixture(scope="session")
def harness_rsq():
    return _harness_lib("libhostcheck_rsq.so")
