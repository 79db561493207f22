"""
Batched replicas (aimd_amd.batch) on the GPU: every replica of a batch against
the single-molecule hf backend at the same geometry (energy 1e-9 Eh, forces
1e-7 Eh/bohr), a small batch against the oracle, and a few Langevin BAOAB
steps of the whole batch on device tensors.
"""

import os

import numpy as np
import pytest

from tests.conftest import ANG, ROOT

pytestmark = pytest.mark.gpu

KINDS = {"water": ("water.xyz", 0, 1), "methanol": ("methanol.xyz", 0, 1), "h4": ("h4_cluster.xyz", 0, 1),
         "methyl": ("methyl_radical.xyz", 0, 2)}


def _replicas(names, seed=3, sigma_ang=0.05):
    from aimd_amd.md import read_xyz
    rng = np.random.default_rng(seed)
    mols, xs = [], []
    for nm in names:
        f, c, m = KINDS[nm]
        s, x = read_xyz(os.path.join(ROOT, "tests", "golden", "replicas", f))
        mols.append((s, c, m))
        xs.append((x + rng.normal(0.0, sigma_ang, x.shape)) * ANG)
    return mols, xs


def test_batch_matches_single_molecule_backend():
    from aimd_amd.backends import get_backend
    from aimd_amd.batch import ReplicaBatch
    names = ["water", "methanol", "methyl", "water", "h4", "methanol", "methyl", "water"]
    mols, xs = _replicas(names)
    batch = ReplicaBatch(mols, basis="6-31g**", conv_tol=1e-11, conv_tol_grad=1e-9)
    out = batch.compute(np.concatenate(xs))
    print("batch iterations", out["iterations"], "converged", out["converged"], flush=True)
    assert out["converged"].all()
    off = 0
    for r, ((s, c, m), x) in enumerate(zip(mols, xs)):
        be = get_backend("hf")(s, charge=c, multiplicity=m, basis="6-31g**", conv_tol=1e-11, conv_tol_grad=1e-9)
        ref = be.compute(x)
        g = out["gradient"][off:off + len(s)]
        de, dg = abs(out["energy"][r] - ref.energy), np.abs(g - ref.gradient).max()
        print(f"replica {r} {names[r]}: dE {de:.2e} dG {dg:.2e}", flush=True)
        assert de < 1e-9 and dg < 1e-7
        off += len(s)
    # a second geometry reuses the densities of the first as guesses
    out2 = batch.compute(np.concatenate(xs) + 0.01)
    assert out2["converged"].all() and out2["iterations"] < out["iterations"]
    batch.close()


def test_batch_against_oracle(oracle_lib):
    from aimd_amd.batch import ReplicaBatch
    names = ["water", "methyl", "h4"]
    mols, xs = _replicas(names, seed=9)
    batch = ReplicaBatch(mols, basis="6-31g**", conv_tol=1e-11, conv_tol_grad=1e-9)
    out = batch.compute(np.concatenate(xs))
    off = 0
    for r, ((s, c, m), x) in enumerate(zip(mols, xs)):
        st, g = oracle_lib.energy_and_gradient(s, x, "6-31g**", mult=m, conv_energy=1e-11, conv_commutator=1e-9)
        de = abs(out["energy"][r] - st.energy)
        dg = np.abs(out["gradient"][off:off + len(s)] - g).max()
        print(f"replica {r} {names[r]} vs oracle: dE {de:.2e} dG {dg:.2e}", flush=True)
        assert de < 1e-9 and dg < 1e-7
        off += len(s)
    batch.close()


def test_batch_baoab_without_friction_is_velocity_verlet_per_replica():
    """BAOAB at zero friction and temperature reduces to velocity Verlet: every replica of
    the batch follows the single-molecule VelocityVerlet trajectory of the hf backend."""
    import torch
    from aimd_amd import md
    from aimd_amd.backends import get_backend
    from aimd_amd.batch import BatchLangevinBAOAB, ReplicaBatch
    names = ["water", "methanol", "methyl", "water"]
    mols, xs = _replicas(names, seed=5)
    kw = dict(conv_tol=1e-11, conv_tol_grad=1e-9)
    batch = ReplicaBatch(mols, basis="6-31g**", **kw)
    v = [md.maxwell_boltzmann(md.masses_au(s), 300.0, seed=r) for r, (s, _, _) in enumerate(mols)]
    integ = BatchLangevinBAOAB(batch, 0.5, 0.0, friction_per_fs=0.0, seed=1)
    dev = batch.device
    integ.initialize(torch.tensor(np.concatenate(xs), device=dev), torch.tensor(np.concatenate(v), device=dev))
    singles = []
    for (s, c, m), x, vr in zip(mols, xs, v):
        be = get_backend("hf")(s, charge=c, multiplicity=m, basis="6-31g**", **kw)
        sysm = md.System.create(s, x, vr)
        vv = md.VelocityVerlet(be, 0.5)
        vv.initialize(sysm)
        singles.append((sysm, vv))
    for step in range(3):
        r = integ.step()
        assert r["converged"].all()
        xb = integ.x.cpu().numpy()
        for k, ((sysm, vv), a, b) in enumerate(zip(singles, batch.rep_atom[:-1], batch.rep_atom[1:])):
            vv.step(sysm)
            dx = np.abs(xb[a:b] - sysm.positions).max()
            de = abs(r["energy"][k] - vv.result.energy)
            assert dx < 1e-8 and de < 1e-9, (step, k, dx, de)
    batch.close()


def test_batch_baoab_thermostat_bookkeeping():
    """At 300 K with friction every replica stays converged, and E_pot + E_kin - heat moves
    only by the integration error (the O step's kinetic-energy change is booked as heat)."""
    import torch
    from aimd_amd.batch import BatchLangevinBAOAB, ReplicaBatch
    from aimd_amd.md import maxwell_boltzmann
    names = ["water", "methanol", "methyl"] * 4
    mols, xs = _replicas(names, seed=5, sigma_ang=0.01)
    batch = ReplicaBatch(mols, basis="6-31g**")
    v = np.concatenate([maxwell_boltzmann(batch.masses[a:b], 300.0, seed=r)
                        for r, (a, b) in enumerate(zip(batch.rep_atom[:-1], batch.rep_atom[1:]))])
    integ = BatchLangevinBAOAB(batch, 0.5, 300.0, friction_per_fs=0.01, seed=1)
    dev = batch.device
    integ.initialize(torch.tensor(np.concatenate(xs), device=dev), torch.tensor(v, device=dev))
    e0 = integ.conserved_energy()
    for _ in range(5):
        assert integ.step()["converged"].all()
    assert float(integ.heat.abs().max()) > 0.0
    drift = np.abs(integ.conserved_energy() - e0).max()
    print("batch BAOAB: 5 steps, max |effective energy change|", drift, flush=True)
    assert drift < 1e-3
    batch.close()
