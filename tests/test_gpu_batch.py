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


def test_batch_baoab_steps():
    import torch
    from aimd_amd.batch import BatchLangevinBAOAB, ReplicaBatch
    from aimd_amd.md import maxwell_boltzmann
    names = ["water", "methanol", "h4", "methyl"] * 4
    mols, xs = _replicas(names, seed=5)
    batch = ReplicaBatch(mols, basis="6-31g**")
    v = np.concatenate([maxwell_boltzmann(batch.masses[a:b], 300.0, seed=r)
                        for r, (a, b) in enumerate(zip(batch.rep_atom[:-1], batch.rep_atom[1:]))])
    md = BatchLangevinBAOAB(batch, 0.5, 300.0, friction_per_fs=0.001, seed=1)
    dev = batch.device
    md.initialize(torch.tensor(np.concatenate(xs), device=dev), torch.tensor(v, device=dev))
    e0 = md.conserved_energy()
    for _ in range(5):
        r = md.step()
        assert r["converged"].all()
    drift = np.abs(md.conserved_energy() - e0).max()
    print("batch BAOAB: 5 steps, max |effective energy change|", drift, flush=True)
    assert drift < 1e-4
    batch.close()
