"""
Bases with f and g shells on the GPU (the runtime-l kernels): every integral
of the reference's f/g case (He with f, H with g, contracted,
tests/test_qc_integrals.py:322-365 of the reference) against the oracle, whose
f/g path tests/test_oracle.py pins; J/K bitwise reproducible; and an SCF
energy and gradient with an f shell on O against the oracle.
"""

import numpy as np
import pytest

from tests.conftest import WATER, bohr
from tests.test_oracle import FG_DATA, FG_POS

pytestmark = pytest.mark.gpu


def _pair(oracle_lib, symbols, pos, table):
    from aimd_amd.qc.basis import basis_from_table
    b = basis_from_table(symbols, pos, table)
    ob = oracle_lib.make_basis(symbols, pos, table)
    assert b.nao == ob.nao and np.array_equal(b.shell_l, ob.shell_l)
    return b, ob


def test_fg_integrals_against_oracle(oracle_lib):
    import torch
    from aimd_amd.qc.engine import Engine
    O = oracle_lib
    b, ob = _pair(O, ["He", "H"], FG_POS, FG_DATA)
    assert (b.lmax, b.nao) == (4, 36)
    e = Engine(b)
    e.set_geometry(FG_POS, 0.0)
    dev = e.device
    n = b.nao
    S, T, V = (torch.empty((n, n), dtype=torch.float64, device=dev) for _ in range(3))
    M = torch.empty((3, n, n), dtype=torch.float64, device=dev)
    e.one_electron(S, T, V, M)
    So, To, Vo, Mo = O.one_electron(ob)
    for got, ref, name in ((S, So, "S"), (T, To, "T"), (V, Vo, "V"), (M, Mo, "M")):
        err = np.abs(got.cpu().numpy() - ref).max()
        print(f"{name}: max |GPU - oracle| {err:.2e}", flush=True)
        assert err < 1e-12
    rng = np.random.default_rng(17)
    A = rng.standard_normal((2, n, n)) * 0.1
    Ds = np.einsum("sij,skj->sik", A, A)
    J, K = e.jk(torch.tensor(Ds, device=dev))
    Jo, Ko = O.Pairs(ob).jk_direct(Ds, 0.0)
    assert np.abs(J.cpu().numpy() - Jo).max() < 1e-11 and np.abs(K.cpu().numpy() - Ko).max() < 1e-11
    J2, K2 = e.jk(torch.tensor(Ds, device=dev))
    assert torch.equal(J, J2) and torch.equal(K, K2)
    D = Ds[0]
    W = rng.standard_normal((n, n)) * 0.1
    W = W + W.T
    kin, nuc, ovl = e.grad_one_terms(torch.tensor(D, device=dev), torch.tensor(W, device=dev))
    gk, gv, gs = O.one_electron_gradient_terms(ob, D, W)
    for got, ref in ((kin, gk), (nuc, gv), (ovl, gs)):
        assert np.abs(got.cpu().numpy() - ref).max() < 1e-11
    g1 = e.grad_one(torch.tensor(D, device=dev), torch.tensor(W, device=dev)).cpu().numpy()
    assert np.abs(g1 - (gk + gv + gs)).max() < 1e-11
    Dt = torch.tensor(D, device=dev)
    g2 = e.grad_two(Dt, Dt, 0.5, threshold=0.0).cpu().numpy()
    g2o = O.Pairs(ob).eri_gradient(D, [D], 0.5, thr=0.0)
    print(f"ERI gradient: max |GPU - oracle| {np.abs(g2 - g2o).max():.2e}", flush=True)
    assert np.abs(g2 - g2o).max() < 1e-11
    assert np.abs(g2.sum(axis=0)).max() < 1e-12
    gu = e.grad_two(torch.tensor(Ds.sum(0), device=dev), torch.tensor(Ds, device=dev), 1.0, threshold=0.0).cpu().numpy()
    guo = O.Pairs(ob).eri_gradient(Ds.sum(0), [Ds[0], Ds[1]], 1.0, thr=0.0)
    assert np.abs(gu - guo).max() < 1e-11
    e.close()


def _water_f_table():
    from aimd_amd.qc.basis_data import BASIS_SETS
    t = BASIS_SETS["6-31g*"]
    return {"O": list(t["O"]) + [[3, [0.8, 1.0]]], "H": list(t["H"])}


@pytest.mark.parametrize("mult", [1, 3])
def test_scf_with_an_f_shell_against_oracle(oracle_lib, mult):
    from aimd_amd.qc.basis import basis_from_table
    from aimd_amd.qc.gradients import scf_gradient
    from aimd_amd.qc.scf import SCFSolver
    s, x = bohr(WATER, 0.03, 5)
    table = _water_f_table()
    b = basis_from_table(s, x, table)
    solver = SCFSolver(s, basis=b, multiplicity=mult, conv_energy=1e-11, conv_commutator=1e-9)
    res = solver.run(x)
    g = scf_gradient(res)
    st, go = oracle_lib.energy_and_gradient(s, x, table, mult=mult, conv_energy=1e-11, conv_commutator=1e-9)
    de, dg = abs(res.energy - st.energy), np.abs(np.asarray(g) - go).max()
    print(f"water + f (mult {mult}): dE {de:.2e} Eh, dG {dg:.2e} Eh/bohr", flush=True)
    assert res.converged and de < 1e-9 and dg < 1e-7
    solver.close()
