"""
GPU parity of the engine's building blocks against the oracle, called through
the C ABI (aimd_amd.qc.engine): Schwarz bounds and the screened quartet list
(bit-exact set), J/K (1e-11), one-electron matrices, both gradient terms,
bitwise run-to-run reproducibility.
"""

import numpy as np
import pytest

from tests.conftest import METHYL, WATER, bohr, cluster, water_dimer

pytestmark = pytest.mark.gpu


def _engine(s, x, basis, mode="auto"):
    import torch
    from aimd_amd.qc.basis import build_basis
    from aimd_amd.qc.engine import Engine
    b = build_basis(s, x, basis)
    e = Engine(b)
    e.set_jk_mode(mode)
    e.set_geometry(x)
    torch.cuda.synchronize()
    return b, e


CASES = [("water", "6-31g*"), ("water", "cc-pvdz"), ("water", "sto-3g"), ("dimer", "6-31g**"),
         ("water8", "6-31g*")]


def _geom(name):
    if name == "dimer":
        return water_dimer()
    if name == "water8":
        return cluster("water8.xyz")
    return bohr(WATER, 0.05, 9)


@pytest.mark.parametrize("mode", ["direct", "store"])
@pytest.mark.parametrize("mol,basis", CASES)
def test_schwarz_quartets_and_jk(oracle_lib, mol, basis, mode):
    import torch
    s, x = _geom(mol)
    print(f"case {mol} {basis} {mode}", flush=True)
    b, e = _engine(s, x, basis, mode)
    assert e.jk_info()["stored"] == (mode == "store")
    ob = oracle_lib.make_basis(s, x, basis)
    pr = oracle_lib.Pairs(ob)
    Qo = pr.schwarz()
    np.testing.assert_allclose(e.schwarz_bounds(), Qo, rtol=1e-14, atol=0)
    ql = e.quartets()
    if b.nblock == b.nshell:
        oq = pr.quartets(1e-12)
        assert len(ql) == len(oq)
        key = lambda a: a[:, 0].astype(np.int64) * (1 << 32) + a[:, 1]
        np.testing.assert_array_equal(np.sort(key(ql)), np.sort(key(oq)))
        # (general contractions and the threshold's edge: test_gpu_parity.py)
    rng = np.random.default_rng(3)
    D = rng.standard_normal((2, b.nao, b.nao))
    D = D + D.transpose(0, 2, 1)
    J, K = e.jk(torch.tensor(D, device="cuda"))
    J2, K2 = e.jk(torch.tensor(D, device="cuda"))
    assert torch.equal(J, J2) and torch.equal(K, K2)
    Jo, Ko = pr.jk_direct(D)
    np.testing.assert_allclose(J.cpu().numpy(), Jo, rtol=0, atol=1e-11)
    np.testing.assert_allclose(K.cpu().numpy(), Ko, rtol=0, atol=1e-11)
    e.close()


@pytest.mark.parametrize("mol,basis,mult", [("water", "6-31g*", 1), ("water", "cc-pvdz", 1), ("methyl", "6-31g*", 2),
                                            ("water8", "6-31g*", 1)])
def test_one_electron_and_gradient_terms(oracle_lib, mol, basis, mult):
    import torch
    s, x = bohr(METHYL, 0.03, 1) if mol == "methyl" else _geom(mol)
    print(f"case {mol} {basis} {mult}", flush=True)
    b, e = _engine(s, x, basis)
    n = b.nao
    S = torch.empty((n, n), dtype=torch.float64, device="cuda")
    T, V = torch.empty_like(S), torch.empty_like(S)
    M = torch.empty((3, n, n), dtype=torch.float64, device="cuda")
    e.one_electron(S, T, V, M)
    ob = oracle_lib.make_basis(s, x, basis)
    So, To, Vo, Mo = oracle_lib.one_electron(ob)
    for a, o in ((S, So), (T, To), (V, Vo), (M, Mo)):
        np.testing.assert_allclose(a.cpu().numpy(), o, rtol=0, atol=1e-12)
    st = oracle_lib.scf(ob, mult=mult, conv_commutator=1e-9)
    spin = np.array([st.Ds[0], st.Ds[-1]])
    Fs = np.array([st.F[0], st.F[-1]])
    P = spin[0] + spin[1]
    W = spin[0] @ Fs[0] @ spin[0] + spin[1] @ Fs[1] @ spin[1]
    W = 0.5 * (W + W.T)
    g1 = e.grad_one(torch.tensor(P, device="cuda"), torch.tensor(W, device="cuda")).cpu().numpy()
    np.testing.assert_allclose(g1, oracle_lib.one_electron_gradient(ob, P, W), rtol=0, atol=1e-11)
    terms = e.grad_one_terms(torch.tensor(P, device="cuda"), torch.tensor(W, device="cuda"))
    for got, want in zip(terms, oracle_lib.one_electron_gradient_terms(ob, P, W)):
        np.testing.assert_allclose(got.cpu().numpy(), want, rtol=0, atol=1e-11)
    Dx, k = (P[None], 0.5) if mult == 1 else (spin, 1.0)
    g2 = e.grad_two(torch.tensor(P, device="cuda"), torch.tensor(Dx, device="cuda"), k).cpu().numpy()
    g2b = e.grad_two(torch.tensor(P, device="cuda"), torch.tensor(Dx, device="cuda"), k).cpu().numpy()
    np.testing.assert_array_equal(g2, g2b)
    g2o = oracle_lib.Pairs(ob).eri_gradient(P, list(Dx), k)
    np.testing.assert_allclose(g2, g2o, rtol=0, atol=1e-11)
    e.close()


@pytest.mark.parametrize("mode,precision", [("store", "fp64"), ("store", "mixed"), ("direct", "fp64")])
def test_density_difference_builds(oracle_lib, mode, precision):
    """J/K of a density difference (incremental Fock): linear against the oracle
    with nothing screened, exactly zero when every unit is screened out, and
    FP32 digestion in mixed precision within FP32 rounding of the FP64 result."""
    import torch
    from aimd_amd.qc.basis import build_basis
    from aimd_amd.qc.engine import Engine
    s, x = _geom("water8")
    print(f"case water8 6-31g* {mode} {precision}", flush=True)
    b = build_basis(s, x, "6-31g*")
    e = Engine(b)
    e.set_jk_mode(mode)
    e.set_precision(precision, 1e-5, 1e-7, 1.0)
    e.set_geometry(x)
    rng = np.random.default_rng(11)
    dD = rng.standard_normal((1, b.nao, b.nao)) * 1e-4
    dD = dD + dD.transpose(0, 2, 1)
    scale = np.abs(dD).max()
    dt = torch.tensor(dD, device="cuda")
    dJ, dK = e.jk_delta(dt, 0.0)
    Jo, Ko = oracle_lib.Pairs(oracle_lib.make_basis(s, x, "6-31g*")).jk_direct(dD)
    # FP64: the oracle's 1e-11 at unit density, scaled; FP32 digestion: ~1e-7 relative per term
    tol = (2e-5 if precision == "mixed" else 1e-10) * scale
    np.testing.assert_allclose(dJ.cpu().numpy(), Jo, rtol=0, atol=tol)
    np.testing.assert_allclose(dK.cpu().numpy(), Ko, rtol=0, atol=tol)
    dJ2, dK2 = e.jk_delta(dt, 1e-300)
    assert torch.equal(dJ, dJ2) and torch.equal(dK, dK2)
    if mode == "store":
        # every Schwarz product is <= ~1e2, so units * max|dD| stay far below 1e3
        z = e.jk_delta(dt, 1e3)
        assert not torch.any(z[0]) and not torch.any(z[1])
        # screening at 1e-9 skips units with q q < 1e-9 / max|dD|: each skipped
        # integral-density product is below 1e-9 (Schwarz), and an element sums
        # far fewer than 1e3 of them at full size
        dJs, dKs = e.jk_delta(dt, 1e-9)
        assert not (torch.equal(dJs, dJ) and torch.equal(dKs, dK))
        np.testing.assert_allclose(dJs.cpu().numpy(), Jo, rtol=0, atol=tol + 1e3 * 1e-9)
        np.testing.assert_allclose(dKs.cpu().numpy(), Ko, rtol=0, atol=tol + 1e3 * 1e-9)
    e.close()


def test_unsupported_angular_momentum_is_a_value_error():
    from aimd_amd.qc.basis import BasisSet
    from aimd_amd.qc.engine import Engine
    b = BasisSet("custom", ["H"], np.zeros((1, 3)), [0], [5], [0, 1], [1.0], [1.0])
    with pytest.raises(ValueError, match="l = 5"):
        Engine(b)
