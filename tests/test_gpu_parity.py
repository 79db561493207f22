"""
GPU parity of the pieces the energy and force checks cover only indirectly,
each against the oracle through the C ABI:

* the screened quartet list of general-contraction bases (cc-pVDZ), by
  expanding the oracle's block quartets into shell quartets, with the cases at
  the threshold's edge counted and decided the same way;
* Schwarz bounds to a few ulp, kept primitive pairs per shell pair;
* FP32 derivative integrals (every quartet on the FP32 path) at fixed densities;
* the gradient's screened list: a superset of the reference rule whose extra
  quartets all lie below the threshold;
* the SAD guess density; the mixed path at (H2O)16 against the oracle fixture;
* an SCF from a non-idempotent guess (what XL-BOMD hands over).
"""

import json
import os

import numpy as np
import pytest

from tests.conftest import HCL, METHYL, ROOT, WATER, bohr, cluster, water_dimer

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


def _geom(name):
    if name == "dimer":
        return water_dimer()
    if name == "water8":
        return cluster("water8.xyz")
    if name == "hcl":
        return bohr(HCL, 0.02, 5)
    return bohr(WATER, 0.05, 9)


def _shell_pair_blocks(b):
    """Block pair index of every shell pair s1 >= s2 (index s1 (s1 + 1) / 2 + s2)."""
    s1, s2 = np.tril_indices(b.nshell)
    b1, b2 = b.shell_block[s1], b.shell_block[s2]
    hi, lo = np.maximum(b1, b2), np.minimum(b1, b2)
    order = np.argsort(s1 * (s1 + 1) // 2 + s2)
    return (hi * (hi + 1) // 2 + lo)[order]


def _expand_block_quartets(b, oq):
    """Shell quartets (P >= Q) of the oracle's block quartets (p1 >= p2)."""
    bp = _shell_pair_blocks(b)
    members = {}
    for P, B in enumerate(bp.tolist()):
        members.setdefault(B, []).append(P)
    out = []
    for p1, p2 in oq.tolist():
        A, C = members.get(p1, []), members.get(p2, [])
        for P in A:
            for Q in C:
                if p1 == p2 and Q > P:
                    continue
                out.append((max(P, Q), min(P, Q)))
    return np.array(out, np.int64).reshape(-1, 2)


def _key(a):
    return a[:, 0].astype(np.int64) * (1 << 32) + a[:, 1]


@pytest.mark.parametrize("mol,basis", [("water", "cc-pvdz"), ("hcl", "cc-pvdz"), ("water8", "cc-pvdz"),
                                       ("dimer", "6-31g**"), ("water8", "6-31g*")])
def test_quartet_list_schwarz_and_pairs_against_oracle(oracle_lib, mol, basis):
    s, x = _geom(mol)
    print(f"case {mol} {basis}", flush=True)
    b, e = _engine(s, x, basis)
    ob = oracle_lib.make_basis(s, x, basis)
    pr = oracle_lib.Pairs(ob)
    pair_pp, pair_blocks, pp_ab, _, Qo = pr.arrays()
    Qg = e.schwarz_bounds()
    ulp = np.abs(Qg - Qo) / np.spacing(np.maximum(np.abs(Qo), 1e-300))
    print(f"  Schwarz: max {ulp.max():.1f} ulp over {len(Qo)} block pairs", flush=True)
    assert ulp.max() <= 8
    thr = 1e-12
    oq = pr.quartets(thr)
    want = _expand_block_quartets(b, oq)
    got = e.quartets()
    assert len(got) == len(want)
    np.testing.assert_array_equal(np.sort(_key(got)), np.sort(_key(want)))
    # the threshold's edge: every block quartet whose product lies within 1e-9 relative
    # of the threshold is decided by the GPU as the oracle decides it
    p1, p2 = np.tril_indices(len(Qo))
    prod = Qo[p1] * Qo[p2]
    edge = np.abs(prod / thr - 1.0) < 1e-9
    if edge.any():
        gset = set(_key(got).tolist())
        for a, c in zip(p1[edge].tolist(), p2[edge].tolist()):
            exp = _expand_block_quartets(b, np.array([[a, c]]))
            inside = all(k in gset for k in _key(exp).tolist()) if len(exp) else True
            assert inside == (Qo[a] * Qo[c] >= thr)
    print(f"  {len(got)} shell quartets, {int(edge.sum())} block products at the threshold's edge", flush=True)
    # kept primitive pairs of every shell pair = the oracle's kept exponent pairs of its
    # block pair restricted to the two shells' exponents (same mu r^2 <= 50 rule)
    npp = e.pair_npp()
    bp = _shell_pair_blocks(b)
    s1, s2 = np.tril_indices(b.nshell)
    order = np.argsort(s1 * (s1 + 1) // 2 + s2)
    s1, s2 = s1[order], s2[order]
    for P in range(len(npp)):
        B = bp[P]
        i, j = pair_blocks[B]
        kept = pp_ab[pair_pp[B]:pair_pp[B + 1]]
        sa, sb = (s1[P], s2[P]) if b.shell_block[s1[P]] == i else (s2[P], s1[P])
        ea = set(b.prim_exp[b.shell_prim[sa]:b.shell_prim[sa + 1]].tolist())
        eb = set(b.prim_exp[b.shell_prim[sb]:b.shell_prim[sb + 1]].tolist())
        n = sum(1 for a_, b_ in kept.tolist() if a_ in ea and b_ in eb)
        assert npp[P] == n, (P, sa, sb, npp[P], n)
    e.close()


@pytest.mark.parametrize("case", ["water8_rhf", "methyl_uhf"])
def test_fp32_derivative_integrals_against_oracle(oracle_lib, case):
    """Every quartet above the gradient threshold on the FP32 derivative path (the switch
    set above every bound), contracted with converged oracle densities. FP32 keeps about
    7 digits per contribution; the gradient sums contributions whose magnitudes add up to
    O(10-100) Eh/bohr per atom, so the expected error is O(1e-6) Eh/bohr."""
    import torch
    from aimd_amd.qc.basis import build_basis
    from aimd_amd.qc.engine import Engine
    if case == "water8_rhf":
        (s, x), mult = cluster("water8.xyz"), 1
    else:
        (s, x), mult = bohr(METHYL, 0.03, 1), 2
    print(f"case {case}", flush=True)
    ob = oracle_lib.make_basis(s, x, "6-31g*")
    st = oracle_lib.scf(ob, mult=mult, conv_commutator=1e-9, stored=False)
    spin = np.array([st.Ds[0], st.Ds[-1]])
    P = spin[0] + spin[1]
    Dx, k = (P[None], 0.5) if mult == 1 else (spin, 1.0)
    g_ref = oracle_lib.Pairs(ob).eri_gradient(P, list(Dx), k)
    b = build_basis(s, x, "6-31g*")
    e = Engine(b)
    e.set_precision("mixed", 1e-5, 1e30, 1e-2)
    e.set_geometry(x)
    g32 = e.grad_two(torch.tensor(P, device="cuda"), torch.tensor(Dx, device="cuda"), k).cpu().numpy()
    e.set_precision("fp64")
    g64 = e.grad_two(torch.tensor(P, device="cuda"), torch.tensor(Dx, device="cuda"), k).cpu().numpy()
    err32, err64 = np.abs(g32 - g_ref).max(), np.abs(g64 - g_ref).max()
    print(f"  max |g_fp32 - oracle| {err32:.2e}, fp64 {err64:.2e}, max |g| {np.abs(g_ref).max():.2f}", flush=True)
    assert err64 < 1e-11
    assert 0.0 < err32 < 2e-5
    e.close()


def test_gradient_list_is_a_superset_of_the_reference_rule(oracle_lib):
    """The GPU pre-pass bounds max|Gamma| from per-pair density maxima: it must keep
    every quartet the reference keeps (Q Q max|Gamma| >= thr, not one-centre), and every
    quartet it keeps beyond those must have Q Q max|Gamma| below thr."""
    import torch
    s, x = cluster("water8.xyz")
    b, e = _engine(s, x, "6-31g*")
    assert b.nblock == b.nshell
    ob = oracle_lib.make_basis(s, x, "6-31g*")
    st = oracle_lib.scf(ob, conv_commutator=1e-8, stored=False)
    P = 2 * st.Ds[0]
    for thr in (1e-14, 1e-10):
        e.grad_two(torch.tensor(P, device="cuda"), torch.tensor(P[None], device="cuda"), 0.5, thr)
        got = e.grad_quartets()
        pr = oracle_lib.Pairs(ob)
        n = pr.npair
        p1, p2 = np.tril_indices(n)
        cand = np.stack([p1, p2], 1).astype(np.int64)
        bound = pr.grad_bound(P, [P], 0.5, cand)
        ref = cand[bound >= thr]
        gk = set(_key(got).tolist())
        rk = set(_key(ref).tolist())
        assert rk <= gk, len(rk - gk)
        extra = np.array(sorted(gk - rk), np.int64)
        if len(extra):
            eb = pr.grad_bound(P, [P], 0.5, np.stack([extra >> 32, extra & 0xFFFFFFFF], 1))
            assert np.all(eb < thr), eb.max()
        print(f"  thr {thr:g}: reference {len(rk)}, GPU {len(gk)} ({len(extra)} extra, all below)", flush=True)
    e.close()


@pytest.mark.parametrize("mol,basis,mult", [("water", "6-31g*", 1), ("methyl", "6-31g*", 2),
                                            ("water", "cc-pvdz", 1), ("hcl", "cc-pvdz", 1)])
def test_sad_guess_against_oracle(oracle_lib, mol, basis, mult):
    from aimd_amd.qc.basis import build_basis
    from aimd_amd.qc.scf import sad_density
    s, x = bohr(METHYL, 0.03, 1) if mol == "methyl" else _geom(mol)
    P = sad_density(build_basis(s, x, basis))
    Po = oracle_lib.sad_density(oracle_lib.make_basis(s, x, basis))
    d = np.abs(P - Po).max()
    print(f"case {mol} {basis}: max |P_sad - oracle| {d:.2e}", flush=True)
    assert d < 1e-6


def test_water16_mixed_against_oracle_fixture():
    """The metric's bounds against the oracle itself: |dE| <= 1 uEh, |dF| <= 1e-5 Eh/bohr."""
    from aimd_amd.backends import get_backend
    with open(os.path.join(ROOT, "tests", "golden", "oracle_water16_631gs.json")) as f:
        ref = json.load(f)
    s, x = cluster("water16.xyz")
    be = get_backend("hf")(s, basis="6-31g*", conv_tol=1e-11, conv_tol_grad=1e-9,
                           scf_options={"precision": "mixed"})
    r = be.compute(x)
    de = abs(r.energy - ref["energy"])
    df = np.abs(r.gradient - np.array(ref["gradient"])).max()
    print(f"(H2O)16 mixed vs oracle: dE {de:.2e} Eh, max dF {df:.2e} Eh/bohr", flush=True)
    assert r.converged and de < 1e-6 and df < 1e-5


def test_non_idempotent_guess_converges():
    """Reference tests/test_qc_scf.py:384-392: XL-BOMD hands over auxiliary densities
    that are not idempotent."""
    from aimd_amd.qc.scf import SCFSolver
    s, x = bohr(WATER)
    solver = SCFSolver(s, "6-31g*")
    r0 = solver.run(x)
    r1 = solver.run(x + np.random.default_rng(3).normal(0.0, 0.05, x.shape))
    A = np.random.default_rng(0).standard_normal((solver.nao, solver.nao))
    mix = 0.6 * r0.density + 0.4 * r1.density + 1e-3 * (A + A.T)
    r = solver.run(x, guess=mix)
    assert r.converged and abs(r.energy - r0.energy) < 1e-9
