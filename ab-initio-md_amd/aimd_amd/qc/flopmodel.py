"""
Algorithmic work model of the integral-direct Fock build and the
derivative-ERI gradient, counted on the Schwarz-surviving quartets (FMA = 2
flops; exp, sqrt, divide = 1 op each).

For a quartet of class (l_a l_b | l_c l_d), L = l_a+l_b+l_c+l_d, m_P / m_K the
kept primitive pairs of bra / ket, n_cd the ket Cartesian count and
box(ab) = prod_d (a_d + b_d + 1):

    per primitive quartet  F_q = (20 + 3L) + 12 + 3 sum_{m=1..L} (NH(m) - 1)
                                 + NH(l_a+l_b) NH(l_c+l_d) (1 + 2 n_cd)
    per bra primitive pair F_b = 2 n_cd sum_ab box(ab)
    per quartet            F   = m_P m_K F_q + m_P F_b
    digestion per contracted integral: 4 (J) + 8 (K) flops per density set.

The gradient (grad_work) counts each pass (bra centres, ket centres) with
L+1, NH(l_a+l_b+1) and 6 derivative FMAs per element of the extended Hermite
box; a pass is counted when the derivative kernel runs it (a pair on one atom
is covered by translational invariance).
Byte model: 8 bytes per primitive-pair record double read once per quartet.
"""

from __future__ import annotations

import numpy as np

from aimd_amd.qc.basis import BasisSet


def NH(L: int) -> int:
    return (L + 1) * (L + 2) * (L + 3) // 6 if L >= 0 else 0


def NC(l: int) -> int:
    return (l + 1) * (l + 2) // 2


def _comps(l):
    return [(lx, ly, l - lx - ly) for lx in range(l, -1, -1) for ly in range(l - lx, -1, -1)]


def _box_sum(la, lb, extra=0):
    s = 0
    for a in _comps(la):
        for b in _comps(lb):
            s += (a[0] + b[0] + 1 + extra) * (a[1] + b[1] + 1 + extra) * (a[2] + b[2] + 1 + extra)
    return s


PAIR_L = [(0, 0), (1, 0), (1, 1), (2, 0), (2, 1), (2, 2)]
CLASS_PAIRS = [(hi, lo) for hi in range(6) for lo in range(hi + 1)]   # class k = hi(hi+1)/2 + lo


def class_ls(k: int) -> tuple[int, int, int, int]:
    hi, lo = CLASS_PAIRS[k]
    return PAIR_L[hi] + PAIR_L[lo]


def per_prim_quartet(la, lb, lc, ld, deriv=0) -> float:
    L = la + lb + lc + ld + deriv
    ncd = NC(lc) * NC(ld)
    r = sum(NH(m) - 1 for m in range(1, L + 1))
    return (20 + 3 * L) + 12 + 3 * r + NH(la + lb + deriv) * NH(lc + ld) * (1 + 2 * ncd)


def per_bra_prim(la, lb, lc, ld, deriv=0) -> float:
    ncd = NC(lc) * NC(ld)
    f = 2 * ncd * _box_sum(la, lb, deriv)
    return 6 * f if deriv else f


def pair_classes(basis: BasisSet) -> np.ndarray:
    """Pair class (ss, ps, pp, ds, dp, dd = 0..5) of every shell pair s1 >= s2."""
    l = basis.shell_l
    s1, s2 = np.tril_indices(basis.nshell)
    o = np.lexsort((s2, s1))
    s1, s2 = s1[o], s2[o]
    la = np.maximum(l[s1], l[s2])
    lb = np.minimum(l[s1], l[s2])
    return la * (la + 1) // 2 + lb


def jk_work(basis: BasisSet, quartets: np.ndarray, npp: np.ndarray, nset: int = 1) -> dict:
    """Per-class and total algorithmic flops of one J/K build over ``quartets`` ((n, 2) pair indices)."""
    pc = pair_classes(basis)
    P, Q = quartets[:, 0], quartets[:, 1]
    cp, cq = pc[P], pc[Q]
    bra = np.where(cp >= cq, P, Q)
    ket = np.where(cp >= cq, Q, P)
    hi, lo = np.maximum(cp, cq), np.minimum(cp, cq)
    k = hi * (hi + 1) // 2 + lo
    mP = npp[bra].astype(np.float64)
    mK = npp[ket].astype(np.float64)
    out = {"per_class": np.zeros(21), "quartets": np.bincount(k, minlength=21),
           "prim_quartets": np.bincount(k, weights=mP * mK, minlength=21), "integrals": np.zeros(21)}
    sum_mm = out["prim_quartets"]
    sum_mp = np.bincount(k, weights=mP, minlength=21)
    for c in range(21):
        la, lb, lc, ld = class_ls(c)
        nint = NC(la) * NC(lb) * NC(lc) * NC(ld)
        out["integrals"][c] = out["quartets"][c] * nint
        out["per_class"][c] = (sum_mm[c] * per_prim_quartet(la, lb, lc, ld) + sum_mp[c] * per_bra_prim(la, lb, lc, ld)
                               + out["quartets"][c] * nint * (4 + 8) * nset)
    out["total"] = float(out["per_class"].sum())
    return out


def grad_work(basis: BasisSet, quartets: np.ndarray, npp: np.ndarray) -> dict:
    """Per-class and total algorithmic flops of one derivative-ERI contraction over
    ``quartets`` ((n, 2) pair indices, the quartets that pass the Gamma screening)."""
    pc = pair_classes(basis)
    l = basis.shell_l
    s1, s2 = np.tril_indices(basis.nshell)
    o = np.lexsort((s2, s1))
    s1, s2 = s1[o], s2[o]
    atom = basis.shell_atom
    one_centre = atom[s1] == atom[s2]                 # per pair: both shells on one atom
    P, Q = quartets[:, 0], quartets[:, 1]
    cp, cq = pc[P], pc[Q]
    bra = np.where(cp >= cq, P, Q)
    ket = np.where(cp >= cq, Q, P)
    hi, lo = np.maximum(cp, cq), np.minimum(cp, cq)
    k = hi * (hi + 1) // 2 + lo
    mP = npp[bra].astype(np.float64)
    mK = npp[ket].astype(np.float64)
    need_b = ~one_centre[bra]
    need_k = ~one_centre[ket]
    out = {"per_class": np.zeros(21), "quartets": np.bincount(k, minlength=21)}
    for c in range(21):
        sel = k == c
        if not sel.any():
            continue
        la, lb, lc, ld = class_ls(c)
        nab, ncd = NC(la) * NC(lb), NC(lc) * NC(ld)
        nb, nk = need_b[sel], need_k[sel]
        run_b = nb | (~nk & (nab >= ncd))
        run_k = nk | (~nb & (nab < ncd))
        mp, mk = mP[sel], mK[sel]
        f = 0.0
        f += np.sum(run_b * (mp * mk * per_prim_quartet(la, lb, lc, ld, 1) + mp * per_bra_prim(la, lb, lc, ld, 1)))
        f += np.sum(run_k * (mk * mp * per_prim_quartet(lc, ld, la, lb, 1) + mk * per_bra_prim(lc, ld, la, lb, 1)))
        out["per_class"][c] = f
    out["total"] = float(out["per_class"].sum())
    return out

