"""
MD with the GPU backend: XL-BOMD (device-resident auxiliary density) against
BOMD at a loose SCF tolerance, bit-for-bit continuation through a checkpoint
file, NVE trajectories and energy drift against the oracle-driven fixtures
((H2O)8/cc-pVDZ 20 steps, (H2O)16/6-31G* 5 steps), 100 steps of (H2O)8/cc-pVDZ,
50 steps of (H2O)16/6-31G*,
and the (H2O)32/6-31G* single point (FP64 and mixed) against the oracle.
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
    return get_backend("hf")(s, basis=basis, **kw)


def test_xlbomd_history_on_device_and_less_drift_than_bomd_at_loose_scf():
    """Reference aimd/xlbomd.py:1-60: with a loosely converged SCF, the time-reversible
    auxiliary-density guess removes the systematic drift that previous-density guesses
    cause. Water/6-31G*, 0.5 fs, 400 K, SCF 1e-5 Eh / 1e-3, 120 steps."""
    import torch
    from aimd_amd import md
    s, x = bohr(WATER)
    v = md.maxwell_boltzmann(md.masses_au(s), 400.0, seed=3)
    drift = {}
    for kind in ("BOMD", "XL"):
        be = _hf(s, "6-31g*", conv_tol=1e-5, conv_tol_grad=1e-3)
        sysm = md.System.create(s, x, v)
        integ = md.XLBOMD(be, 0.5, k=5) if kind == "XL" else md.VelocityVerlet(be, 0.5)
        integ.initialize(sysm)
        if kind == "XL":
            assert isinstance(integ.aux.history, torch.Tensor) and integ.aux.history.is_cuda
        e = [integ.conserved_energy(sysm)]
        for _ in range(120):
            integ.step(sysm)
            e.append(integ.conserved_energy(sysm))
        t = np.arange(len(e))
        drift[kind] = abs(np.polyfit(t, np.array(e), 1)[0])
    print("straight-line drift per step (Eh): BOMD", drift["BOMD"], "XL-BOMD", drift["XL"], flush=True)
    assert drift["XL"] < drift["BOMD"]


@pytest.mark.parametrize("kind", ["VelocityVerlet", "XLBOMD"])
def test_checkpoint_restart_is_bit_for_bit_on_the_gpu(tmp_path, kind):
    from aimd_amd import md
    s, x = bohr(WATER, 0.02, 4)
    v = md.maxwell_boltzmann(md.masses_au(s), 300.0, seed=2)

    def make(be):
        return md.XLBOMD(be, 0.5) if kind == "XLBOMD" else md.VelocityVerlet(be, 0.5)

    kw = dict(conv_tol=1e-9, conv_tol_grad=1e-6)
    ref_sys = md.System.create(s, x, v)
    ref = make(_hf(s, "6-31g*", **kw))
    ref.initialize(ref_sys)
    for _ in range(8):
        ref.step(ref_sys)
    run_sys = md.System.create(s, x, v)
    run = make(_hf(s, "6-31g*", **kw))
    run.initialize(run_sys)
    for _ in range(4):
        run.step(run_sys)
    ck = md.load_checkpoint(md.save_checkpoint(tmp_path / "ck.npz", run_sys, run, 4))
    be2 = _hf(s, "6-31g*", **kw)
    ck.check_backend(be2)
    cont_sys = md.System.create(s, x, v)
    cont = make(be2)
    ck.restore(cont_sys, cont)
    for _ in range(4):
        cont.step(cont_sys)
    np.testing.assert_array_equal(cont_sys.positions, ref_sys.positions)
    np.testing.assert_array_equal(cont_sys.velocities, ref_sys.velocities)
    assert cont.result.energy == ref.result.energy


@pytest.mark.parametrize("name,xyz,basis,steps", [("nve_water8", "water8.xyz", "cc-pvdz", 20),
                                                  ("nve_water16", "water16.xyz", "6-31g*", 5)])
def test_nve_trajectory_against_the_oracle(name, xyz, basis, steps):
    from aimd_amd.md import VelocityVerletNVE, masses_au, maxwell_boltzmann
    fx = os.path.join(ROOT, "tests", "golden", f"oracle_{name}_{basis.replace('-', '').replace('*', 's')}.json")
    if not os.path.exists(fx):
        pytest.skip(f"fixture {os.path.basename(fx)} not generated")
    ref = _fixture(os.path.basename(fx))
    s, x = cluster(xyz)
    m = masses_au(s)
    v = maxwell_boltzmann(m, 300.0, seed=1)
    traj = VelocityVerletNVE(_hf(s, basis, conv_tol=1e-10, conv_tol_grad=1e-8), m, x, v, 0.5)
    traj.initialize()
    e_tot = [traj.total_energy()]
    for k in range(steps):
        traj.step()
        e_tot.append(traj.total_energy())
        dx = np.abs(traj.x - np.array(ref["positions"][k + 1])).max()
        assert dx < 1e-6, (k, dx)
        assert abs(traj.result.energy - ref["energy"][k + 1]) < 1e-7
    span, span_o = np.ptp(e_tot), np.ptp(ref["total_energy"][:steps + 1])
    print(f"{name}: total-energy span GPU {span:.3e}, oracle {span_o:.3e}", flush=True)
    assert span <= span_o * 1.01 + 1e-8


def test_water8_ccpvdz_100_nve_steps():
    """Energy conservation of (H2O)8/cc-pVDZ over 100 steps at 0.5 fs."""
    from aimd_amd.md import VelocityVerletNVE, masses_au, maxwell_boltzmann
    s, x = cluster("water8.xyz")
    m = masses_au(s)
    traj = VelocityVerletNVE(_hf(s, "cc-pvdz", conv_tol=1e-10, conv_tol_grad=1e-8), m, x,
                             maxwell_boltzmann(m, 300.0, seed=1), 0.5)
    traj.initialize()
    e = [traj.total_energy()]
    for _ in range(100):
        traj.step()
        e.append(traj.total_energy())
    e = np.array(e)
    slope = np.polyfit(np.arange(len(e)), e, 1)[0]
    print(f"(H2O)8/cc-pVDZ 100 steps: span {np.ptp(e):.3e} Eh, drift {slope * 2000:.3e} Eh/ps", flush=True)
    assert np.ptp(e) < 1e-3


def test_water16_631gs_50_nve_steps():
    """The (H2O)16/6-31G* trajectory of the benchmark: 50 steps at 0.5 fs from 300 K, the
    first 5 on the oracle-driven trajectory, total energy bounded with no systematic drift."""
    from aimd_amd.md import VelocityVerletNVE, masses_au, maxwell_boltzmann
    ref = _fixture("oracle_nve_water16_631gs.json")
    s, x = cluster("water16.xyz")
    m = masses_au(s)
    traj = VelocityVerletNVE(_hf(s, "6-31g*", conv_tol=1e-10, conv_tol_grad=1e-8), m, x,
                             maxwell_boltzmann(m, 300.0, seed=1), 0.5)
    traj.initialize()
    e = [traj.total_energy()]
    for k in range(50):
        traj.step()
        e.append(traj.total_energy())
        if k < len(ref["energy"]) - 1:
            assert np.abs(traj.x - np.array(ref["positions"][k + 1])).max() < 1e-6
            assert abs(traj.result.energy - ref["energy"][k + 1]) < 1e-7
    e = np.array(e)
    slope = np.polyfit(np.arange(len(e)), e, 1)[0]
    print(f"(H2O)16/6-31G* 50 steps: span {np.ptp(e):.3e} Eh, straight-line change {slope * 50:.3e} Eh", flush=True)
    assert np.ptp(e) < 3e-3 and abs(slope * 50) < 1e-3


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_water32_single_point_against_oracle(precision):
    ref = _fixture("oracle_water32_631gs.json")
    s, x = cluster("water32.xyz")
    r = _hf(s, "6-31g*", conv_tol=1e-11, conv_tol_grad=1e-9, scf_options={"precision": precision}).compute(x)
    de = abs(r.energy - ref["energy"])
    dg = np.abs(r.gradient - np.array(ref["gradient"])).max()
    print(f"(H2O)32 {precision}: dE {de:.2e} Eh, max dF {dg:.2e} Eh/bohr", flush=True)
    if precision == "fp64":
        assert de < 1e-9 and dg < 1e-7
    else:
        assert de < 1e-6 and dg < 1e-5
