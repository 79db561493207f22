"""
Full-size workloads on the GPU against the oracle's committed fixtures
(tests/golden/oracle_*.json, from make_oracle_fixtures.py) and through
size-independent properties: (H2O)16/6-31G* (304 AOs), (H2O)8/cc-pVDZ
(200 AOs, forces also against central differences), determinism,
translation/rotation invariance, and the 20-step water/6-31G* NVE trajectory
against the oracle's.
"""

import json
import os

import numpy as np
import pytest

from tests.conftest import ROOT, WATER, bohr, cluster

pytestmark = pytest.mark.gpu


def _fixture(name):
    with open(os.path.join(ROOT, "tests", "golden", name)) as f:
        return json.load(f)


def _hf(s, basis, **kw):
    from aimd_amd.backends import get_backend
    return get_backend("hf")(s, basis=basis, **{**dict(conv_tol=1e-11, conv_tol_grad=1e-9), **kw})


@pytest.mark.parametrize("jk_mode", ["store", "direct"])
def test_water16_631gs_matches_oracle_and_is_deterministic(jk_mode):
    s, x = cluster("water16.xyz")
    ref = _fixture("oracle_water16_631gs.json")
    be = _hf(s, "6-31g*", scf_options={"jk_mode": jk_mode})
    r = be.compute(x)
    assert be.solver.engine.jk_info()["stored"] == (jk_mode == "store")
    print("(H2O)16", jk_mode, r.energy, ref["energy"], r.info["scf_iterations"], flush=True)
    assert abs(r.energy - ref["energy"]) < 1e-9
    assert np.abs(r.gradient - np.array(ref["gradient"])).max() < 1e-7
    # translation invariance (net force) and bitwise reproducibility
    assert np.abs(r.gradient.sum(axis=0)).max() < 1e-9
    guess = be.last_scf.density.copy()
    be.set_density_guess(guess)
    r2 = be.compute(x)
    be.set_density_guess(guess)
    r3 = be.compute(x)
    assert r2.energy == r3.energy
    np.testing.assert_array_equal(r2.gradient, r3.gradient)


def test_water16_rotation_covariance():
    from scipy.spatial.transform import Rotation
    s, x = cluster("water16.xyz")
    be = _hf(s, "6-31g*", conv_tol=1e-10, conv_tol_grad=1e-8)
    r = be.compute(x)
    R = Rotation.from_rotvec([0.3, -0.2, 0.5]).as_matrix()
    c = x.mean(axis=0)
    xr = (x - c) @ R.T + c
    be.reset_guess()
    rr = be.compute(xr)
    assert rr.energy == pytest.approx(r.energy, abs=1e-8)
    np.testing.assert_allclose(rr.gradient, r.gradient @ R.T, rtol=0, atol=2e-7)


def test_water8_ccpvdz_forces_against_oracle_and_central_differences():
    s, x = cluster("water8.xyz")
    ref = _fixture("oracle_water8_ccpvdz.json")
    be = _hf(s, "cc-pvdz", conv_tol=1e-12, conv_tol_grad=1e-10)
    r = be.compute(x)
    assert abs(r.energy - ref["energy"]) < 1e-9
    assert np.abs(r.gradient - np.array(ref["gradient"])).max() < 1e-7
    rng = np.random.default_rng(1)
    h = 1e-3
    for idx in rng.choice(x.size, 12, replace=False):
        a, d = divmod(int(idx), 3)
        y1, y2 = x.copy(), x.copy()
        y1[a, d] -= h
        y2[a, d] += h
        fd = (be.compute(y2).energy - be.compute(y1).energy) / (2 * h)
        assert r.gradient[a, d] == pytest.approx(fd, abs=2e-6), (a, d)


def test_water_631gs_nve_trajectory_matches_oracle(oracle_lib):
    """Config 1: 20 velocity-Verlet steps at 0.5 fs from 300 K velocities (seed 1), GPU vs oracle."""
    from aimd_amd.md import VelocityVerletNVE, masses_au, maxwell_boltzmann
    O = oracle_lib
    s, x = bohr(WATER)
    m = masses_au(s)
    v = maxwell_boltzmann(m, 300.0, seed=1)
    be = _hf(s, "6-31g*", conv_tol=1e-10, conv_tol_grad=1e-8)
    md = VelocityVerletNVE(be, m, x, v, 0.5)
    md.initialize()

    class OracleBackend:
        def __init__(self):
            self.P = None

        def compute(self, y):
            from types import SimpleNamespace
            st, g = O.energy_and_gradient(s, y, "6-31g*", conv_energy=1e-10, conv_commutator=1e-8,
                                          guess=self.P if self.P is not None else "sad")
            self.P = 2 * st.Ds[0]
            return SimpleNamespace(energy=st.energy, gradient=g)

    mo = VelocityVerletNVE(OracleBackend(), m, x, v, 0.5)
    mo.initialize()
    e0, eo0 = md.total_energy(), mo.total_energy()
    dev, devo = 0.0, 0.0
    for _ in range(20):
        md.step()
        mo.step()
        dev = max(dev, abs(md.total_energy() - e0))
        devo = max(devo, abs(mo.total_energy() - eo0))
    assert np.abs(md.x - mo.x).max() < 1e-7
    assert abs(md.result.energy - mo.result.energy) < 1e-8
    assert dev <= devo * 1.01 + 1e-9


def test_water16_mixed_precision_within_the_metric_bounds():
    # mixed: FP32 for quartets below the switches; the metric's bounds are |dE| <= 1 uEh and
    # force error <= 1e-5 Eh/bohr against the FP64 path, and results stay bitwise reproducible
    s, x = cluster("water16.xyz")
    r64 = _hf(s, "6-31g*").compute(x)
    bem = _hf(s, "6-31g*", scf_options={"precision": "mixed"})
    rm = bem.compute(x)
    de, df = abs(rm.energy - r64.energy), np.abs(rm.gradient - r64.gradient).max()
    print("(H2O)16 mixed vs fp64: dE", de, "max dF", df, flush=True)
    assert de < 1e-6
    assert df < 1e-5
    guess = bem.last_scf.density.copy()
    bem.set_density_guess(guess)
    r2 = bem.compute(x)
    bem.set_density_guess(guess)
    r3 = bem.compute(x)
    assert r2.energy == r3.energy
    np.testing.assert_array_equal(r2.gradient, r3.gradient)
