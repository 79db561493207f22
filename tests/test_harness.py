"""
The gfx950 kernels' arithmetic, run on the CPU through the check harness
(tests/harness, the same __host__ __device__ templates) against the oracle:
J/K, Schwarz bounds, the screened quartet set, one-electron integrals and
both gradient terms. No GPU needed.
"""

import numpy as np
import pytest

from tests.conftest import METHYL, WATER, bohr, water_dimer

CASES = [("water", "sto-3g"), ("water", "6-31g*"), ("water", "6-31g**"), ("water", "cc-pvdz"),
         ("dimer", "6-31g*")]


def _mol(name):
    if name == "dimer":
        return water_dimer()
    return bohr(WATER, 0.08, 4)


@pytest.mark.parametrize("mol,basis", CASES)
def test_jk_schwarz_and_quartets_match_oracle(harness, oracle_lib, mol, basis):
    from aimd_amd.qc.basis import build_basis
    s, x = _mol(mol)
    b = build_basis(s, x, basis)
    ob = oracle_lib.make_basis(s, x, basis)
    rng = np.random.default_rng(1)
    D = rng.standard_normal((2, b.nao, b.nao))
    D = D + D.transpose(0, 2, 1)
    J, K, qs, qb, qu = harness.jk(b, D)
    pr = oracle_lib.Pairs(ob)
    Jo, Ko = pr.jk_direct(D)
    np.testing.assert_allclose(J, Jo, rtol=0, atol=1e-12)
    np.testing.assert_allclose(K, Ko, rtol=0, atol=1e-12)
    Qo = pr.schwarz()
    np.testing.assert_allclose(qb, Qo, rtol=4e-15, atol=0)
    if b.nblock == b.nshell:            # segmented basis: shell quartets are block quartets
        assert {tuple(q) for q in qu} == {tuple(q) for q in pr.quartets(1e-12)}


@pytest.mark.parametrize("mol,basis,mult", [("water", "6-31g*", 1), ("water", "cc-pvdz", 1),
                                            ("methyl", "6-31g*", 2)])
def test_one_electron_and_gradients_match_oracle(harness, oracle_lib, mol, basis, mult):
    from aimd_amd.qc.basis import build_basis
    s, x = bohr(METHYL, 0.03, 1) if mol == "methyl" else _mol(mol)
    b = build_basis(s, x, basis)
    ob = oracle_lib.make_basis(s, x, basis)
    st = oracle_lib.scf(ob, mult=mult, conv_commutator=1e-9)
    spin = np.array([st.Ds[0], st.Ds[-1]])
    Fs = np.array([st.F[0], st.F[-1]])
    P = spin[0] + spin[1]
    W = spin[0] @ Fs[0] @ spin[0] + spin[1] @ Fs[1] @ spin[1]
    W = 0.5 * (W + W.T)
    S, T, V, M, g1 = harness.one_electron(b, P, W)
    So, To, Vo, Mo = oracle_lib.one_electron(ob)
    for a, o in ((S, So), (T, To), (V, Vo), (M, Mo)):
        np.testing.assert_allclose(a, o, rtol=0, atol=1e-13)
    np.testing.assert_allclose(g1, oracle_lib.one_electron_gradient(ob, P, W), rtol=0, atol=1e-12)
    if mult == 1:
        Dx, k = P[None], 0.5
    else:
        Dx, k = spin, 1.0
    g2 = harness.grad_two(b, P, Dx, k)
    g2o = oracle_lib.Pairs(ob).eri_gradient(P, list(Dx), k)
    np.testing.assert_allclose(g2, g2o, rtol=0, atol=1e-12)


@pytest.mark.parametrize("mol,basis,mult", [("water", "6-31g*", 1), ("water", "cc-pvdz", 1),
                                            ("methyl", "6-31g*", 2), ("dimer", "6-31g**", 1)])
def test_device_derivative_work_items_match_oracle(harness, harness_rsq, oracle_lib, mol, basis, mult):
    """The gfx950 derivative kernel's own work items -- column groups per pass, the
    staged Boys slice with its exp(-T_k) column -- run on the host in FP64 agree with
    the oracle at 1e-11; in FP32 within FP32 rounding; and with a model of the
    hardware reciprocal-square-root estimate (relative error 5e-8 before the Newton
    steps), used in both the prefactor and the large-T Boys branch, still at 1e-11:
    the device-only force error of that variant is not in this arithmetic."""
    from aimd_amd.qc.basis import build_basis
    s, x = bohr(METHYL, 0.03, 1) if mol == "methyl" else _mol(mol)
    b = build_basis(s, x, basis)
    ob = oracle_lib.make_basis(s, x, basis)
    st = oracle_lib.scf(ob, mult=mult, conv_commutator=1e-9)
    spin = np.array([st.Ds[0], st.Ds[-1]])
    P = spin[0] + spin[1]
    Dx, k = (P[None], 0.5) if mult == 1 else (spin, 1.0)
    g2o = oracle_lib.Pairs(ob).eri_gradient(P, list(Dx), k)
    g64 = harness.grad_two_variant(1, b, P, Dx, k)
    np.testing.assert_allclose(g64, g2o, rtol=0, atol=1e-11)
    g32 = harness.grad_two_variant(2, b, P, Dx, k)
    err32 = np.abs(g32 - g2o).max()
    assert 0.0 < err32 < 1e-5, err32
    grsq = harness_rsq.grad_two_variant(1, b, P, Dx, k)
    np.testing.assert_allclose(grsq, g2o, rtol=0, atol=1e-11)
