"""
Gaussian basis-set tables for H-Ar: STO-3G, 6-31G, 6-31G*, 6-31G**, cc-pVDZ.

The numbers live in ``basis_data.json`` next to this file, in the row layout
the reference uses (``aimd/qc/basis_data.py:1-12``):

    {basis: {element: [[l, [exp_1, c_11, c_12, ...], [exp_2, ...], ...], ...]}}

one row per primitive (exponent in bohr^-2, then its coefficient in every
contracted function of the shell; coefficients refer to normalised
primitives). The values are the reference's (which equal PySCF 2.14's
copies), so energies are comparable digit for digit.
"""

from __future__ import annotations

import json
import os

import numpy as np

from aimd_amd.elements import normalize_symbol

_PATH = os.path.join(os.path.dirname(os.path.abspath(__file__)), "basis_data.json")
with open(_PATH, "r", encoding="utf-8") as _f:
    BASIS_SETS: dict[str, dict[str, list]] = json.load(_f)

_ALIASES = {
    "sto3g": "sto-3g",
    "631g": "6-31g",
    "6-31g(d)": "6-31g*",
    "631g*": "6-31g*",
    "6-31g(d,p)": "6-31g**",
    "631g**": "6-31g**",
    "ccpvdz": "cc-pvdz",
}


def available_basis_sets() -> list[str]:
    return list(BASIS_SETS)


def normalize_basis_name(name: str) -> str:
    """Canonical lower-case name: '6-31G*' -> '6-31g*', '6-31G(d)' -> '6-31g*'
    (``aimd/qc/basis_data.py:1051-1059``)."""
    key = "".join(str(name).split()).lower()
    key = _ALIASES.get(key, key)
    if key not in BASIS_SETS:
        raise ValueError(
            f"Unknown basis set {name!r}. Available: {', '.join(available_basis_sets())}")
    return key


def get_basis_shells(name: str, symbol: str) -> list[tuple[int, np.ndarray, np.ndarray]]:
    """
    Shells of ``symbol`` in basis ``name`` as (l, exponents (nprim,),
    coefficients (nprim, ncontr)) in table order (``aimd/qc/basis_data.py:1066-1085``).
    """
    key = normalize_basis_name(name)
    sym = normalize_symbol(symbol)
    table = BASIS_SETS[key]
    if sym not in table:
        raise ValueError(f"Basis set {key!r} has no data for element {sym!r} "
                         f"(available: {', '.join(table)})")
    out = []
    for entry in table[sym]:
        rows = np.array(entry[1:], dtype=float)
        out.append((int(entry[0]), rows[:, 0].copy(), rows[:, 1:].copy()))
    return out
