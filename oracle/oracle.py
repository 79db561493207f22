"""
CPU oracle of the Hartree-Fock energy-and-forces path.

TEST INFRASTRUCTURE ONLY. Only ``tests/``, ``__graft_entry__.smoke()`` and the
``cpu_baseline`` leg of ``bench.py`` import this module, and only to check or
to time against; the product (``ab-initio-md_amd/``) never imports it.

It restates the reference's algorithm:

* basis layout               aimd/qc/basis.py:85-122, 185-228, 291-312
* integrals (C++, liboracle)  aimd/qc/boys.py, hermite.py, integrals.py
* SCF driver (numpy)          aimd/qc/scf.py:284-303, 341-511, 516-568, 863-965
* gradient assembly           aimd/qc/gradients.py:106-125, 168-193
* nuclear repulsion           aimd/qc/integrals.py:1421-1449

Checks that pin it (tests/test_oracle.py): the reference's own known answers
(Boys vs 60-digit series; water/STO-3G -74.965901192299 Eh; H2/STO-3G
-1.1167 Eh; CH3 UHF/6-31G* -39.558992 Eh, <S^2> 0.7615; CH3OH/6-31G*
-115.035418 Eh; stretched-H2 RHF/UHF energies; basis layout), closed forms
(s-function overlap/ERI, F_n(0)), invariances (net force, torque, rotation
covariance), central finite differences of the energy, and agreement of the
stored-tensor and integral-direct J/K paths.
"""

from __future__ import annotations

import ctypes
import json
import math
import os
import subprocess
from dataclasses import dataclass, field

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
# AIMD_ORACLE_LIB selects another build of the same source (the sanitizer build)
_LIB_PATH = os.environ.get("AIMD_ORACLE_LIB") or os.path.join(HERE, "liboracle.so")
_DATA = os.path.join(os.path.dirname(HERE), "ab-initio-md_amd", "aimd_amd", "qc", "basis_data.json")

_lib = None


def build(force: bool = False) -> str:
    """Compile liboracle.so with make (g++, OpenMP)."""
    if force or not os.path.exists(_LIB_PATH):
        subprocess.run(["make", "-C", HERE, "liboracle.so"] + (["-B"] if force else []),
                       check=True, stdout=subprocess.DEVNULL)
    return _LIB_PATH


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        L = ctypes.CDLL(_LIB_PATH)
        P = ctypes.c_void_p
        d, i, l = ctypes.c_double, ctypes.c_int, ctypes.c_long
        L.oracle_threads.argtypes = [i]
        L.oracle_threads.restype = i
        L.oracle_boys.argtypes = [i, d, P]
        L.oracle_boys_table.argtypes = [P]
        L.oracle_hermite_e.argtypes = [i, i, d, d, d, d, P]
        L.oracle_hermite_r.argtypes = [i, d, d, d, d, P]
        L.oracle_one_electron.argtypes = [i, P, P, P, P, P, P, P, i, i, P, P, P, P, P, P, P]
        L.oracle_one_electron_grad.argtypes = [i, P, P, P, P, P, P, P, i, i, P, P, P, P]
        L.oracle_one_electron_grad_terms.argtypes = [i, P, P, P, P, P, P, P, i, i, P, P, P, P, P, P]
        L.oracle_pairs_create.argtypes = [i, P, P, P, P, P, P, P, i, i, P]
        L.oracle_pairs_create.restype = P
        L.oracle_pairs_destroy.argtypes = [P]
        L.oracle_pairs_npair.argtypes = [P]
        L.oracle_pairs_npair.restype = i
        L.oracle_pairs_npp.argtypes = [P]
        L.oracle_pairs_npp.restype = l
        L.oracle_pairs_get.argtypes = [P, P, P, P, P, P]
        L.oracle_quartets.argtypes = [P, d, P, l]
        L.oracle_quartets.restype = l
        L.oracle_eri_tensor.argtypes = [P, d, i, P]
        L.oracle_jk_from_eri.argtypes = [i, i, P, P, P, P]
        L.oracle_jk_direct.argtypes = [P, d, i, i, P, P, P]
        L.oracle_eri_grad.argtypes = [P, d, i, i, P, i, P, d, P]
        L.oracle_grad_bound.argtypes = [P, i, P, i, P, d, P, l, P]
        _lib = L
    return _lib


def _p(a: np.ndarray):
    return a.ctypes.data_as(ctypes.c_void_p)


def set_threads(n: int | None = None) -> int:
    return lib().oracle_threads(int(n or 0))


# ------------------------------------------------------------------ basis
with open(_DATA, "r", encoding="utf-8") as _f:
    _BASIS = json.load(_f)
_ALIAS = {"sto3g": "sto-3g", "631g": "6-31g", "6-31g(d)": "6-31g*", "631g*": "6-31g*",
          "6-31g(d,p)": "6-31g**", "631g**": "6-31g**", "ccpvdz": "cc-pvdz"}
SYMBOLS = ("H He Li Be B C N O F Ne Na Mg Al Si P S Cl Ar").split()
Z_OF = {s: i + 1 for i, s in enumerate(SYMBOLS)}


def _dfact(n: int) -> int:
    r = 1
    while n > 1:
        r *= n
        n -= 2
    return r


def _norm_contraction(l, e, c):
    # aimd/qc/basis.py:110-122
    pn = (2.0 * e / np.pi) ** 0.75 * (4.0 * e) ** (0.5 * l) / np.sqrt(_dfact(2 * l - 1))
    cc = c * pn
    p = e[:, None] + e[None, :]
    S = _dfact(2 * l - 1) * np.pi ** 1.5 / ((2.0 * p) ** l * p ** 1.5)
    return cc / np.sqrt(cc @ S @ cc)


@dataclass
class Basis:
    """Flat shell and contraction-block arrays (aimd/qc/basis.py layout)."""
    symbols: list
    positions: np.ndarray
    shell_atom: np.ndarray
    shell_l: np.ndarray
    shell_prim: np.ndarray
    shell_ao: np.ndarray
    prim_exp: np.ndarray
    prim_coef: np.ndarray
    blk_shell: np.ndarray = field(default=None)
    blk_atom: np.ndarray = field(default=None)
    blk_l: np.ndarray = field(default=None)
    blk_ncon: np.ndarray = field(default=None)
    blk_ao: np.ndarray = field(default=None)
    blk_prim: np.ndarray = field(default=None)
    blk_exp: np.ndarray = field(default=None)
    blk_coef: np.ndarray = field(default=None)

    @property
    def nao(self):
        return int(self.shell_ao[-1])

    @property
    def natm(self):
        return len(self.symbols)

    @property
    def charges(self):
        return np.array([float(Z_OF[s]) for s in self.symbols])

    @property
    def ao_atom(self):
        nc = (self.shell_l + 1) * (self.shell_l + 2) // 2
        return np.repeat(self.shell_atom, nc)


def make_basis(symbols, positions, name) -> Basis:
    """Basis of a named set, or of a table {symbol: [[l, [exp, c_1, ...], ...], ...]}
    in the layout of basis_data.json (extra coefficient columns: general contractions)."""
    if isinstance(name, dict):
        table = {k.strip()[:1].upper() + k.strip()[1:].lower(): v for k, v in name.items()}
    else:
        key = "".join(str(name).split()).lower()
        key = _ALIAS.get(key, key)
        table = _BASIS[key]
    sa, sl, sp, pe, pc = [], [], [0], [], []
    for a, s in enumerate(symbols):
        sym = s.strip()[:1].upper() + s.strip()[1:].lower()
        shells = [(int(e[0]), np.array(e[1:], float)) for e in table[sym]]
        for l, rows in sorted(shells, key=lambda t: t[0]):
            for col in range(rows.shape[1] - 1):
                keep = rows[:, 1 + col] != 0.0
                e = rows[keep, 0]
                c = _norm_contraction(l, e, rows[keep, 1 + col])
                sa.append(a)
                sl.append(l)
                pe.extend(e)
                pc.extend(c)
                sp.append(len(pe))
    sl = np.array(sl, np.int64)
    nc = (sl + 1) * (sl + 2) // 2
    b = Basis([s.strip()[:1].upper() + s.strip()[1:].lower() for s in symbols],
              np.ascontiguousarray(np.array(positions, float).reshape(-1, 3)),
              np.array(sa, np.int64), sl, np.array(sp, np.int64),
              np.concatenate([[0], np.cumsum(nc)]).astype(np.int64),
              np.array(pe, float), np.array(pc, float))
    _blocks(b)
    return b


def _blocks(b: Basis) -> None:
    # aimd/qc/basis.py:185-228 (nested exponent sets on one atom, equal l)
    blk_shell, blk_exp, cols = [0], [], []
    union: list = []
    ns = len(b.shell_l)
    for s in range(ns):
        k0, k1 = b.shell_prim[s], b.shell_prim[s + 1]
        col: dict = {}
        for e, c in zip(b.prim_exp[k0:k1].tolist(), b.prim_coef[k0:k1].tolist()):
            col[e] = col.get(e, 0.0) + c
        es = set(col)
        if (s > 0 and b.shell_atom[s] == b.shell_atom[s - 1] and b.shell_l[s] == b.shell_l[s - 1]
                and (es <= set(union) or set(union) <= es)):
            union.extend(e for e in col if e not in union)
            cols[-1].append(col)
        else:
            if s > 0:
                blk_shell.append(s)
                blk_exp.append(union)
            union = list(col)
            cols.append([col])
    blk_shell.append(ns)
    blk_exp.append(union)
    maxcon = max(len(c) for c in cols)
    nprim = sum(len(u) for u in blk_exp)
    coef = np.zeros((nprim, maxcon))
    prim = [0]
    for bi in range(len(cols)):
        for k, e in enumerate(blk_exp[bi]):
            for c, col in enumerate(cols[bi]):
                coef[prim[-1] + k, c] = col.get(e, 0.0)
        prim.append(prim[-1] + len(blk_exp[bi]))
    bs = np.array(blk_shell, np.int64)
    b.blk_shell = bs
    b.blk_atom = b.shell_atom[bs[:-1]].copy()
    b.blk_l = b.shell_l[bs[:-1]].copy()
    b.blk_ncon = np.diff(bs)
    b.blk_ao = b.shell_ao[bs[:-1]].copy()
    b.blk_prim = np.array(prim, np.int64)
    b.blk_exp = np.array([e for u in blk_exp for e in u], float)
    b.blk_coef = np.ascontiguousarray(coef)


def with_positions(b: Basis, positions) -> Basis:
    nb = Basis(**{k: getattr(b, k) for k in b.__dataclass_fields__})
    nb.positions = np.ascontiguousarray(np.array(positions, float).reshape(-1, 3))
    return nb


# ------------------------------------------------------------- integrals
def boys(nmax: int, x: float) -> np.ndarray:
    out = np.zeros(nmax + 1)
    lib().oracle_boys(int(nmax), float(x), _p(out))
    return out


def boys_table() -> np.ndarray:
    out = np.zeros((402, 43))
    lib().oracle_boys_table(_p(out))
    return out


def one_electron(b: Basis, origin=None):
    """(S, T, V, dipole integrals (3,n,n))."""
    n = b.nao
    S, T, V = np.zeros((n, n)), np.zeros((n, n)), np.zeros((n, n))
    M = np.zeros((3, n, n))
    o = np.zeros(3) if origin is None else np.asarray(origin, float)
    q = b.charges
    lib().oracle_one_electron(len(b.shell_l), _p(b.shell_atom), _p(b.shell_l), _p(b.shell_prim),
                              _p(b.shell_ao), _p(b.prim_exp), _p(b.prim_coef), _p(b.positions), n,
                              b.natm, _p(q), _p(b.positions), _p(o), _p(S), _p(T), _p(M), _p(V))
    return S, T, V, M


def one_electron_gradient(b: Basis, D, W):
    g = np.zeros((b.natm, 3))
    Ds = np.ascontiguousarray(0.5 * (D + D.T))
    Ws = np.ascontiguousarray(0.5 * (W + W.T))
    lib().oracle_one_electron_grad(len(b.shell_l), _p(b.shell_atom), _p(b.shell_l), _p(b.shell_prim),
                                   _p(b.shell_ao), _p(b.prim_exp), _p(b.prim_coef), _p(b.positions),
                                   b.nao, b.natm, _p(b.charges), _p(Ds), _p(Ws), _p(g))
    return g


def one_electron_gradient_terms(b: Basis, D, W):
    """(d/dR sum D T, d/dR sum D V, -d/dR sum W S), the split of
    aimd/qc/gradients.py:196-223."""
    gk, gs, gv = np.zeros((b.natm, 3)), np.zeros((b.natm, 3)), np.zeros((b.natm, 3))
    Ds = np.ascontiguousarray(0.5 * (D + D.T))
    Ws = np.ascontiguousarray(0.5 * (W + W.T))
    lib().oracle_one_electron_grad_terms(len(b.shell_l), _p(b.shell_atom), _p(b.shell_l), _p(b.shell_prim),
                                         _p(b.shell_ao), _p(b.prim_exp), _p(b.prim_coef), _p(b.positions),
                                         b.nao, b.natm, _p(b.charges), _p(Ds), _p(Ws), _p(gk), _p(gs), _p(gv))
    return gk, gv, -gs


class Pairs:
    """Block-pair data and Schwarz bounds of one geometry (C++ handle)."""

    def __init__(self, b: Basis):
        self.b = b
        L = lib()
        self.h = L.oracle_pairs_create(len(b.blk_l), _p(b.blk_atom), _p(b.blk_l), _p(b.blk_prim),
                                       _p(b.blk_ncon), _p(b.blk_ao), _p(b.blk_exp), _p(b.blk_coef),
                                       b.blk_coef.shape[1], b.natm, _p(b.positions))
        self.npair = L.oracle_pairs_npair(self.h)
        self.npp = L.oracle_pairs_npp(self.h)

    def __del__(self):
        try:
            lib().oracle_pairs_destroy(self.h)
        except Exception:
            pass

    def arrays(self):
        pair_pp = np.zeros(self.npair + 1, np.int64)
        pair_blocks = np.zeros((self.npair, 2), np.int64)
        pp_ab = np.zeros((self.npp, 2))
        pp_P = np.zeros((self.npp, 3))
        Q = np.zeros(self.npair)
        lib().oracle_pairs_get(self.h, _p(pair_pp), _p(pair_blocks), _p(pp_ab), _p(pp_P), _p(Q))
        return pair_pp, pair_blocks, pp_ab, pp_P, Q

    def schwarz(self):
        return self.arrays()[4]

    def quartets(self, thr=1e-12) -> np.ndarray:
        L = lib()
        n = L.oracle_quartets(self.h, float(thr), None, 0)
        out = np.zeros((n, 2), np.int64)
        L.oracle_quartets(self.h, float(thr), _p(out), n)
        return out

    def eri(self, thr=1e-12) -> np.ndarray:
        n = self.b.nao
        out = np.zeros((n, n, n, n))
        lib().oracle_eri_tensor(self.h, float(thr), n, _p(out))
        return out

    def jk_direct(self, D, thr=1e-12):
        D = np.asarray(D, float)
        single = D.ndim == 2
        D3 = np.ascontiguousarray(D[None] if single else D)
        J, K = np.zeros_like(D3), np.zeros_like(D3)
        lib().oracle_jk_direct(self.h, float(thr), self.b.nao, D3.shape[0], _p(D3), _p(J), _p(K))
        return (J[0], K[0]) if single else (J, K)

    def eri_gradient(self, Dc, Dx_list, kfac, thr=1e-14):
        n = self.b.nao
        Dc = np.ascontiguousarray(0.5 * (Dc + Dc.T))
        Dx = np.ascontiguousarray(np.array([0.5 * (d + d.T) for d in Dx_list]))
        g = np.zeros((self.b.natm, 3))
        lib().oracle_eri_grad(self.h, float(thr), n, self.b.natm, _p(Dc), len(Dx_list), _p(Dx),
                              float(kfac), _p(g))
        return g


    def grad_bound(self, Dc, Dx_list, kfac, quartets):
        """Q_p1 Q_p2 max|Gamma| of each block quartet (p1, p2), the quantity
        _eri_grad_kernel compares with its threshold; -1 for one-centre quartets."""
        n = self.b.nao
        Dc = np.ascontiguousarray(0.5 * (Dc + Dc.T))
        Dx = np.ascontiguousarray(np.array([0.5 * (d + d.T) for d in Dx_list]))
        qs = np.ascontiguousarray(np.asarray(quartets, np.int64).reshape(-1, 2))
        out = np.zeros(len(qs))
        lib().oracle_grad_bound(self.h, n, _p(Dc), len(Dx_list), _p(Dx), float(kfac), _p(qs), len(qs), _p(out))
        return out


def jk_from_eri(eri, D):
    D = np.asarray(D, float)
    single = D.ndim == 2
    D3 = np.ascontiguousarray(D[None] if single else D)
    J, K = np.zeros_like(D3), np.zeros_like(D3)
    lib().oracle_jk_from_eri(eri.shape[0], D3.shape[0], _p(eri), _p(D3), _p(J), _p(K))
    return (J[0], K[0]) if single else (J, K)


def nuclear_repulsion(q, x):
    i, j = np.triu_indices(len(q), k=1)
    return float(np.sum(q[i] * q[j] / np.linalg.norm(x[i] - x[j], axis=1)))


def nuclear_repulsion_gradient(q, x):
    i, j = np.triu_indices(len(q), k=1)
    d = x[i] - x[j]
    r = np.linalg.norm(d, axis=1)
    pair = -(q[i] * q[j] / r ** 3)[:, None] * d
    g = np.zeros_like(x)
    np.add.at(g, i, pair)
    np.add.at(g, j, -pair)
    return g


# ------------------------------------------------------------------- SCF
def orthogonalizer(S, threshold=1e-7):
    s, U = np.linalg.eigh(S)
    keep = s >= threshold
    if keep.all():
        return (U * s ** -0.5) @ U.T, 0
    return U[:, keep] * s[keep] ** -0.5, int((~keep).sum())


_MADELUNG = sorted(((n, l) for n in range(1, 8) for l in range(n)), key=lambda t: (t[0] + t[1], t[0]))


def _l_counts(Z):
    counts = [0, 0, 0, 0]
    left = Z
    for _n, l in _MADELUNG:
        if left <= 0:
            break
        k = min(left, 2 * (2 * l + 1))
        counts[l] += k
        left -= k
    return counts


def _groups(eps, tol):
    out, i = [], 0
    while i < len(eps):
        j = i + 1
        while j < len(eps) and eps[j] - eps[j - 1] < tol:
            j += 1
        out.append((i, j))
        i = j
    return out


def _atomic_occ(eps, lc, tol=1e-6):
    # aimd/qc/scf.py:394-424
    groups = _groups(eps, tol)
    occ = np.zeros(len(eps))
    left = list(lc) + [0] * (5 - len(lc))
    ok = True
    for i, j in groups:
        g = j - i
        if g % 2 == 0 or g > 9:
            ok = False
            break
        l = (g - 1) // 2
        k = min(left[l], 2 * g)
        occ[i:j] = k / g
        left[l] -= k
    if ok and not any(left):
        return occ
    occ[:] = 0.0
    n = float(sum(lc))
    for i, j in groups:
        k = min(n, 2.0 * (j - i))
        occ[i:j] = k / (j - i)
        n -= k
    return occ


class DIIS:
    """aimd/qc/scf.py:516-568."""

    def __init__(self, space=8, max_coef=1e4):
        self.space, self.max_coef = space, max_coef
        self.F, self.E = [], []

    def __len__(self):
        return len(self.F)

    def reset(self):
        self.F.clear()
        self.E.clear()

    def push(self, f, e):
        self.F.append(np.array(f, float))
        self.E.append(np.array(e, float))
        while len(self.F) > self.space:
            self.F.pop(0)
            self.E.pop(0)

    def coefficients(self):
        last = self.E[-1].ravel()
        if len(self.E) == 1:
            return np.ones(1)
        delta = np.stack([e.ravel() - last for e in self.E[:-1]], axis=1)
        x = np.linalg.lstsq(delta, -last, rcond=None)[0]
        return np.append(x, 1.0 - x.sum())

    def extrapolate(self):
        while len(self.F) > 1:
            c = self.coefficients()
            if np.all(np.isfinite(c)) and np.abs(c).max() <= self.max_coef:
                return sum(ci * Fi for ci, Fi in zip(c, self.F))
            self.F.pop(0)
            self.E.pop(0)
        return self.F[-1].copy()


_SAD_CACHE: dict = {}


def atomic_density(b: Basis, atom: int) -> np.ndarray:
    """SAD block of one atom (aimd/qc/scf.py:430-511), cached per element/basis."""
    sh = np.nonzero(b.shell_atom == atom)[0]
    sym = b.symbols[atom]
    prims = np.concatenate([np.arange(b.shell_prim[s], b.shell_prim[s + 1]) for s in sh])
    key = (sym, b.shell_l[sh].tobytes(), b.prim_exp[prims].tobytes(), b.prim_coef[prims].tobytes())
    if key in _SAD_CACHE:
        return _SAD_CACHE[key].copy()
    sl = b.shell_l[sh]
    sp = np.concatenate([[0], np.cumsum(np.diff(b.shell_prim)[sh])]).astype(np.int64)
    nc = (sl + 1) * (sl + 2) // 2
    ab = Basis([sym], np.zeros((1, 3)), np.zeros(len(sh), np.int64), sl.copy(), sp,
               np.concatenate([[0], np.cumsum(nc)]).astype(np.int64),
               b.prim_exp[prims].copy(), b.prim_coef[prims].copy())
    _blocks(ab)
    Z = Z_OF[sym]
    lc = _l_counts(Z)
    S, T, V, _ = one_electron(ab)
    H = T + V
    s, U = np.linalg.eigh(S)
    keep = s >= 1e-7
    X = U[:, keep] * s[keep] ** -0.5
    if Z == 1:
        eps, Cp = np.linalg.eigh(X.T @ H @ X)
        c = X @ Cp[:, 0]
        D = np.outer(c, c)
        _SAD_CACHE[key] = D
        return D.copy()
    eri = Pairs(ab).eri(1e-12)

    def dens(Fo):
        eps, Cp = np.linalg.eigh(Fo)
        C = X @ Cp
        return (C * _atomic_occ(eps, lc)) @ C.T

    D = dens(X.T @ H @ X)
    diis = DIIS(8)
    e_old = np.inf
    for _ in range(200):
        J, K = jk_from_eri(eri, D)
        F = H + J - 0.5 * K
        F = 0.5 * (F + F.T)
        e = 0.5 * float(np.vdot(D, H + F))
        err = X.T @ (F @ D @ S - S @ D @ F) @ X
        if abs(e - e_old) < 1e-11 and np.abs(err).max() < 1e-8:
            break
        e_old = e
        diis.push((X.T @ F @ X)[None], err[None])
        D = dens(diis.extrapolate()[0])
    D = 0.5 * (D + D.T)
    _SAD_CACHE[key] = D
    return D.copy()


def sad_density(b: Basis) -> np.ndarray:
    n = b.nao
    P = np.zeros((n, n))
    ao_atom = b.ao_atom
    for a in range(b.natm):
        idx = np.nonzero(ao_atom == a)[0]
        if len(idx):
            P[idx[0]:idx[-1] + 1, idx[0]:idx[-1] + 1] = atomic_density(b, a)
    return P


@dataclass
class SCFState:
    energy: float
    converged: bool
    iterations: int
    Ds: np.ndarray          # (nspin, n, n)
    F: np.ndarray           # (nspin, n, n)
    e_el: float
    enuc: float
    err: float
    de: float
    S: np.ndarray
    H: np.ndarray
    M: np.ndarray
    history: list


def electron_counts(b: Basis, charge, mult):
    n = int(sum(Z_OF[s] for s in b.symbols)) - charge
    return (n + mult - 1) // 2, (n - mult + 1) // 2


def scf(b: Basis, charge=0, mult=1, unrestricted=None, guess="sad", conv_energy=1e-10,
        conv_commutator=1e-7, max_iter=100, diis_space=8, diis_restart=12, schwarz=1e-12,
        stored=None, pairs=None) -> SCFState:
    """RHF/UHF loop of aimd/qc/scf.py:863-923 (no level shift / damping)."""
    if unrestricted is None:
        unrestricted = mult != 1
    na, nb = electron_counts(b, charge, mult)
    nocc = (na, nb) if unrestricted else (na,)
    nspin = len(nocc)
    g = 2.0 / nspin
    S, T, V, M = one_electron(b)
    H = T + V
    X, _ = orthogonalizer(S)
    enuc = nuclear_repulsion(b.charges, b.positions)
    pr = pairs or Pairs(b)
    if stored is None:
        stored = b.nao <= 120
    eri = pr.eri(schwarz) if stored else None

    def jk(Ds):
        return jk_from_eri(eri, Ds) if stored else pr.jk_direct(Ds, schwarz)

    n = b.nao
    if isinstance(guess, str):
        P = sad_density(b)
        Ds = np.repeat(0.5 * P[None], nspin, axis=0)
    else:
        D = np.array(guess, float)
        if D.shape == (n, n):
            Ds = np.repeat(0.5 * D[None], nspin, axis=0)
        else:
            Ds = D.copy() if unrestricted else 0.5 * (D[0] + D[1])[None]
        Ds = 0.5 * (Ds + Ds.transpose(0, 2, 1))
    diis = DIIS(diis_space)
    e_prev, best, since, hist = np.nan, np.inf, 0, []
    conv = False
    for it in range(1, max_iter + 1):
        J, K = jk(Ds)
        F = H + g * J.sum(axis=0) - K
        F = 0.5 * (F + F.transpose(0, 2, 1))
        e_el = 0.5 * g * float(sum(np.vdot(Ds[s], H + F[s]) for s in range(nspin)))
        FDS = F @ Ds @ S
        err = X.T @ (FDS - FDS.transpose(0, 2, 1)) @ X
        err_norm = float(np.abs(err).max())
        de = e_el - e_prev
        hist.append((e_el + enuc, de, err_norm))
        if abs(de) < conv_energy and err_norm < conv_commutator:
            conv = True
            break
        if it == max_iter:
            break
        Fo = X.T @ F @ X
        if err_norm < best:
            best, since = err_norm, 0
        else:
            since += 1
        if diis_restart and since >= diis_restart:
            diis.reset()
            best, since = err_norm, 0
        diis.push(Fo, err)
        Fo = diis.extrapolate()
        Dn = np.empty_like(Ds)
        for s in range(nspin):
            _, Cp = np.linalg.eigh(Fo[s])
            C = X @ Cp[:, :nocc[s]]
            Dn[s] = C @ C.T
        Ds = Dn
        e_prev = e_el
    return SCFState(e_el + enuc, conv, it, Ds, F, e_el, enuc, err_norm, de, S, H, M, hist)


def scf_gradient(b: Basis, st: SCFState, pairs=None, thr=1e-14):
    """aimd/qc/gradients.py:168-193, W = sum_s D_s F_s D_s."""
    Ds, F = st.Ds, st.F
    spin = np.array([Ds[0], Ds[-1]])
    Fs = np.array([F[0], F[-1]])
    W = spin[0] @ Fs[0] @ spin[0] + spin[1] @ Fs[1] @ spin[1]
    W = 0.5 * (W + W.T)
    P = spin[0] + spin[1]
    g = nuclear_repulsion_gradient(b.charges, b.positions)
    g = g + one_electron_gradient(b, P, W)
    pr = pairs or Pairs(b)
    if Ds.shape[0] == 2:
        g = g + pr.eri_gradient(P, [spin[0], spin[1]], 1.0, thr)
    else:
        g = g + pr.eri_gradient(P, [P], 0.5, thr)
    return g


def energy_and_gradient(symbols, positions, basis, charge=0, mult=1, **kw):
    b = make_basis(symbols, positions, basis)
    pr = Pairs(b)
    st = scf(b, charge, mult, pairs=pr, **kw)
    return st, scf_gradient(b, st, pr)
