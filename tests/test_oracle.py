"""
Pins the CPU oracle (oracle/) before anything is compared against it: the
reference's own known answers, closed forms, finite differences and
invariances (no GPU needed).
"""

import math
from decimal import Decimal, localcontext

import numpy as np
import pytest
from scipy.special import erf

from tests.conftest import ANG, METHANOL, METHYL, WATER, WATER_STO3G_MIN, bohr, water_dimer


# ------------------------------------------------------------------- Boys
def _boys_series(n, x):
    """F_n(x) = exp(-x) sum_k (2x)^k / ((2n+1)(2n+3)...(2n+2k+1)) to 60 digits."""
    with localcontext() as ctx:
        ctx.prec = 60
        X = Decimal(float(x))
        term = Decimal(1) / (2 * n + 1)
        total = term
        k = 1
        while True:
            term = term * 2 * X / (2 * n + 2 * k + 1)
            total += term
            k += 1
            if term < total * Decimal("1e-40"):
                break
        return float((-X).exp() * total)


def _boys_asym(n, x):
    with localcontext() as ctx:
        ctx.prec = 60
        X = Decimal(float(x))
        pi = Decimal("3.14159265358979323846264338327950288419716939937510582097494")
        return float(math.prod(range(1, 2 * n, 2)) * pi.sqrt() / (2 ** (n + 1) * X ** n * X.sqrt()))


def test_boys_against_60_digit_series(oracle_lib):
    O = oracle_lib
    rng = np.random.default_rng(11)
    xs = list(10 ** rng.uniform(-8, math.log10(150.0), 40)) + [0.0, 0.05, 0.049999999, 1.0, 40 - 1e-9, 40.0,
                                                                40 + 1e-9, 100.0]
    worst = 0.0
    for x in xs:
        f = O.boys(20, x)
        for n in range(21):
            worst = max(worst, abs(f[n] / _boys_series(n, x) - 1.0))
    assert worst < 2e-14
    np.testing.assert_allclose(O.boys(32, 0.0), 1.0 / (2 * np.arange(33) + 1), rtol=1e-15)


def test_boys_large_arguments(oracle_lib):
    worst = 0.0
    for x in [1e3, 1e6, 1e10]:
        f = oracle_lib.boys(24, x)
        for n in range(25):
            r = _boys_asym(n, x)
            if r > 1e-300:
                worst = max(worst, abs(f[n] / r - 1.0))
    assert worst < 2e-14


def test_boys_f0_is_erf_closed_form(oracle_lib):
    for x in [1e-6, 0.3, 2.0, 11.7, 39.0, 55.0]:
        ref = 0.5 * math.sqrt(math.pi / x) * erf(math.sqrt(x))
        assert oracle_lib.boys(0, x)[0] == pytest.approx(ref, rel=2e-15)


# ------------------------------------------------------------ closed forms
def _h2_ss_closed_form(O, R):
    """(11|22), (11|11), S12 of H2/STO-3G from contracted s closed forms (erf-based F_0)."""
    b = O.make_basis(["H", "H"], np.array([[0, 0, 0], [0, 0, R]]), "sto-3g")
    e = b.prim_exp[:3]
    c = b.prim_coef[:3]
    A, B = np.zeros(3), np.array([0.0, 0.0, R])

    def F0(t):
        return 1.0 if t < 1e-15 else 0.5 * math.sqrt(math.pi / t) * erf(math.sqrt(t))

    def eri(C1, C2, C3, C4):
        s = 0.0
        for i in range(3):
            for j in range(3):
                p = e[i] + e[j]
                P = (e[i] * C1 + e[j] * C2) / p
                Kab = math.exp(-e[i] * e[j] / p * np.sum((C1 - C2) ** 2))
                for k in range(3):
                    for l in range(3):
                        q = e[k] + e[l]
                        Q = (e[k] * C3 + e[l] * C4) / q
                        Kcd = math.exp(-e[k] * e[l] / q * np.sum((C3 - C4) ** 2))
                        al = p * q / (p + q)
                        s += (c[i] * c[j] * c[k] * c[l] * 2 * math.pi ** 2.5 / (p * q * math.sqrt(p + q))
                              * Kab * Kcd * F0(al * np.sum((P - Q) ** 2)))
        return s

    S12 = sum(c[i] * c[j] * (math.pi / (e[i] + e[j])) ** 1.5 * math.exp(-e[i] * e[j] / (e[i] + e[j]) * R * R)
              for i in range(3) for j in range(3))
    return b, eri(A, A, A, A), eri(A, A, B, B), eri(A, B, A, B), S12


def test_s_functions_closed_form(oracle_lib):
    O = oracle_lib
    b, e1111, e1122, e1212, S12 = _h2_ss_closed_form(O, 1.4)
    eri = O.Pairs(b).eri(0.0)
    assert eri[0, 0, 0, 0] == pytest.approx(e1111, rel=1e-13)
    assert eri[0, 0, 1, 1] == pytest.approx(e1122, rel=1e-13)
    assert eri[0, 1, 0, 1] == pytest.approx(e1212, rel=1e-13)
    S, T, V, M = O.one_electron(b)
    assert S[0, 1] == pytest.approx(S12, rel=1e-13)
    np.testing.assert_allclose(np.diag(S), 1.0, atol=1e-14)


def test_unit_self_overlap_and_symmetry(oracle_lib):
    s, x = bohr(METHANOL, 0.05, 3)
    b = oracle_lib.make_basis(s, x, "cc-pvdz")
    S, T, V, M = oracle_lib.one_electron(b)
    np.testing.assert_allclose(np.diag(S), 1.0, atol=1e-13)
    for A in (S, T, V, M[0], M[1], M[2]):
        np.testing.assert_allclose(A, A.T, rtol=0, atol=1e-13)


def test_eri_permutational_symmetry_and_schwarz(oracle_lib):
    s, x = bohr(WATER, 0.1, 1)
    b = oracle_lib.make_basis(s, x, "6-31g*")
    pr = oracle_lib.Pairs(b)
    eri = pr.eri(0.0)
    for perm in [(1, 0, 2, 3), (0, 1, 3, 2), (2, 3, 0, 1), (3, 2, 1, 0)]:
        np.testing.assert_array_equal(eri, eri.transpose(perm))
    d = np.einsum("ijij->ij", eri)
    bound = np.sqrt(np.abs(d))
    assert np.all(np.abs(eri) <= bound[:, :, None, None] * bound[None, None, :, :] * (1 + 1e-12) + 1e-300)


def test_direct_jk_equals_stored_tensor_jk(oracle_lib):
    s, x = water_dimer()
    b = oracle_lib.make_basis(s, x, "6-31g*")
    pr = oracle_lib.Pairs(b)
    eri = pr.eri(1e-12)
    rng = np.random.default_rng(2)
    D = rng.standard_normal((2, b.nao, b.nao))
    D = D + D.transpose(0, 2, 1)
    J1, K1 = oracle_lib.jk_from_eri(eri, D)
    J2, K2 = pr.jk_direct(D, 1e-12)
    np.testing.assert_allclose(J2, J1, atol=1e-12)
    np.testing.assert_allclose(K2, K1, atol=1e-12)
    # jk_from_eri itself against einsum
    np.testing.assert_allclose(J1[0], np.einsum("mnls,ls->mn", eri, D[0]), atol=1e-12)
    np.testing.assert_allclose(K1[0], np.einsum("mlns,ls->mn", eri, D[0]), atol=1e-12)


def test_quartet_list_is_the_brute_force_screened_set(oracle_lib):
    s, x = water_dimer()
    b = oracle_lib.make_basis(s, x, "6-31g*")
    pr = oracle_lib.Pairs(b)
    Q = pr.schwarz()
    ql = {tuple(q) for q in pr.quartets(1e-12)}
    brute = {(i, j) for i in range(len(Q)) for j in range(i + 1) if not Q[i] * Q[j] < 1e-12}
    assert ql == brute


# ----------------------------------------------------------- known answers
def test_water_sto3g_minimum_energy(oracle_lib):
    """Reference: tests/test_end_to_end.py:70-71 (PySCF cart=True): -74.965901192299 Eh."""
    s, x = bohr(WATER_STO3G_MIN)
    st, g = oracle_lib.energy_and_gradient(s, x, "sto-3g", conv_energy=1e-11, conv_commutator=1e-9)
    assert st.converged
    assert st.energy == pytest.approx(-74.965901192299, abs=1e-10)
    assert np.abs(g).max() < 1.3e-6          # "max |grad| 1.2e-6" at this HF/STO-3G minimum


def test_h2_sto3g_szabo_ostlund(oracle_lib):
    """Reference: tests/test_hf_backend.py:136-146: R = 1.4 bohr, E = -1.1167 Eh."""
    st, _ = oracle_lib.energy_and_gradient(["H", "H"], np.array([[0, 0, 0], [0, 0, 1.4]]), "sto-3g")
    assert abs(st.energy - (-1.1167)) < 5e-5


def test_methyl_radical_uhf(oracle_lib):
    """Reference README.md:496-500: CH3 UHF/6-31G* E = -39.558992 Eh, <S^2> = 0.7615."""
    s, x = bohr(METHYL)
    b = oracle_lib.make_basis(s, x, "6-31g*")
    st = oracle_lib.scf(b, mult=2, conv_commutator=1e-9, conv_energy=1e-11)
    assert st.energy == pytest.approx(-39.558992, abs=1e-6)
    Da, Db = st.Ds
    s2 = 0.75 + 4 - np.vdot(Da @ st.S, (Db @ st.S).T)
    assert s2 == pytest.approx(0.7615, abs=1e-4)


def test_methanol_rhf(oracle_lib):
    """Reference README.md:496-500: CH3OH RHF/6-31G* E = -115.035418 Eh (max |grad| 1.3e-6)."""
    s, x = bohr(METHANOL)
    st, g = oracle_lib.energy_and_gradient(s, x, "6-31g*")
    assert st.energy == pytest.approx(-115.035418, abs=1e-6)
    assert np.abs(g).max() < 2e-6


def test_stretched_h2_rhf_and_broken_symmetry_uhf(oracle_lib):
    """Reference tests/test_hf_backend.py:265-296: R = 4 bohr, RHF -0.76108, UHF (broken) -0.93584."""
    x = np.array([[0.0, 0.0, 0.0], [0.0, 0.0, 4.0]])
    b = oracle_lib.make_basis(["H", "H"], x, "sto-3g")
    rhf = oracle_lib.scf(b, conv_commutator=1e-9, conv_energy=1e-12)
    uhf = oracle_lib.scf(b, mult=1, unrestricted=True, guess=np.stack([np.diag([1.0, 0.0]), np.diag([0.0, 1.0])]),
                         conv_commutator=1e-9, conv_energy=1e-12)
    assert rhf.energy == pytest.approx(-0.76108, abs=1e-5)
    assert uhf.energy == pytest.approx(-0.93584, abs=1e-5)


# ------------------------------------------------------------ gradients
def test_gradient_matches_central_differences(oracle_lib):
    """5-point FD of the energy (reference tests/test_qc_gradients.py:119-129 uses 3e-9)."""
    s, x = bohr(WATER, 0.05, 7)
    O = oracle_lib
    st, g = O.energy_and_gradient(s, x, "6-31g*", conv_energy=1e-12, conv_commutator=1e-10)
    h = 1e-3

    def E(y):
        return O.energy_and_gradient(s, y, "6-31g*", conv_energy=1e-12, conv_commutator=1e-10,
                                     guess=2 * st.Ds[0])[0].energy

    for a, d in [(0, 2), (1, 0), (2, 1)]:
        vals = []
        for k in (-2, -1, 1, 2):
            y = x.copy()
            y[a, d] += k * h
            vals.append(E(y))
        fd = (vals[0] - 8 * vals[1] + 8 * vals[2] - vals[3]) / (12 * h)
        assert g[a, d] == pytest.approx(fd, abs=2e-8)


def test_gradient_translation_and_rotation_invariance(oracle_lib):
    s, x = bohr(METHANOL, 0.05, 5)
    st, g = oracle_lib.energy_and_gradient(s, x, "6-31g*", conv_energy=1e-12, conv_commutator=1e-10)
    assert np.abs(g.sum(axis=0)).max() < 1e-10
    torque = np.cross(x - x.mean(axis=0), g).sum(axis=0)
    assert np.abs(torque).max() < 1e-8


def test_uhf_gradient_fd(oracle_lib):
    s, x = bohr(METHYL, 0.04, 2)
    O = oracle_lib
    st, g = O.energy_and_gradient(s, x, "6-31g*", mult=2, conv_energy=1e-12, conv_commutator=1e-10)
    h = 1e-3
    y1, y2 = x.copy(), x.copy()
    y1[1, 0] -= h
    y2[1, 0] += h
    e1 = O.energy_and_gradient(s, y1, "6-31g*", mult=2, conv_energy=1e-12, conv_commutator=1e-10)[0].energy
    e2 = O.energy_and_gradient(s, y2, "6-31g*", mult=2, conv_energy=1e-12, conv_commutator=1e-10)[0].energy
    assert g[1, 0] == pytest.approx((e2 - e1) / (2 * h), abs=5e-7)


def test_planted_mistake_is_caught_by_known_answer(oracle_lib, monkeypatch):
    """A wrong sign in the exchange term of the oracle's Fock build misses the known answer."""
    O = oracle_lib
    orig = O.jk_from_eri

    def bad(eri, D):
        J, K = orig(eri, D)
        return J, -K

    monkeypatch.setattr(O, "jk_from_eri", bad)
    s, x = bohr(WATER_STO3G_MIN)
    b = O.make_basis(s, x, "sto-3g")
    st = O.scf(b, max_iter=60)
    assert abs(st.energy - (-74.965901192299)) > 1e-3
