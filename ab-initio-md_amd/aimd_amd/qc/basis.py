"""
Contracted Cartesian Gaussian basis laid out as flat arrays for the device.

Conventions (identical to the reference's, ``aimd/qc/basis.py:1-68``, because
AO order and normalisation are part of the backend's output contract --
densities handed back to callers are AO matrices):

* an AO is f(lx,ly,lz) x^lx y^ly z^lz sum_k c_k exp(-a_k r^2) on its atom,
  every Cartesian component (6 d, 10 f) kept and unit-normalised;
* contraction coefficients normalise the x^l component and include the
  primitive normalisation (2a/pi)^(3/4) (4a)^(l/2) / sqrt((2l-1)!!); component
  (lx,ly,lz) carries f = sqrt((2l-1)!! / ((2lx-1)!! (2ly-1)!! (2lz-1)!!));
* AO order: atoms in input order, shells of an atom sorted by l (stable), a
  general contraction split into one segmented shell per column, components
  in PySCF order (xx, xy, xz, yy, yz, zz).

Contraction blocks: consecutive shells on one atom with equal l whose
exponent sets are nested form one block (``aimd/qc/basis.py:185-228``). The
reference evaluates two-electron integrals per block; this engine evaluates
them per segmented shell on the GPU but screens quartets with the block-pair
Schwarz bound, so both keep exactly the same set of block quartets.

Arrays (int64 / float64, read-only):
    shell_atom, shell_l, shell_prim (nshell+1), shell_ao (nshell+1)
    prim_exp, prim_coef
    shell_block[s]       block that shell s belongs to
    blk_shell (nblock+1), blk_atom, blk_l, blk_ncon, blk_ao
    positions (natm, 3) bohr
"""

from __future__ import annotations

import copy
from dataclasses import dataclass
from typing import Sequence

import numpy as np

from aimd_amd.elements import ATOMIC_NUMBERS, normalize_symbol
from aimd_amd.qc.basis_data import get_basis_shells, normalize_basis_name

_L_LABELS = "spdfghi"
LMAX_DEVICE = 2     # highest l the gfx950 kernels are instantiated for (all built-in bases)


def ncart(l: int) -> int:
    return (l + 1) * (l + 2) // 2


def cartesian_components(l: int) -> list[tuple[int, int, int]]:
    """(lx, ly, lz) of a shell in PySCF order."""
    return [(lx, ly, l - lx - ly) for lx in range(l, -1, -1) for ly in range(l - lx, -1, -1)]


def _double_factorial(n: int) -> int:
    out = 1
    while n > 1:
        out *= n
        n -= 2
    return out


def cartesian_factor(lx: int, ly: int, lz: int) -> float:
    l = lx + ly + lz
    num = _double_factorial(2 * l - 1)
    den = _double_factorial(2 * lx - 1) * _double_factorial(2 * ly - 1) * _double_factorial(2 * lz - 1)
    return float(np.sqrt(num / den))


def normalized_contraction(l: int, exponents: np.ndarray, coefficients: np.ndarray) -> np.ndarray:
    """Coefficients making x^l sum_k c_k exp(-a_k r^2) unit-normalised."""
    a = np.asarray(exponents, dtype=float)
    d = np.asarray(coefficients, dtype=float)
    dfl = _double_factorial(2 * l - 1)
    prim_norm = (2.0 * a / np.pi) ** 0.75 * (4.0 * a) ** (0.5 * l) / np.sqrt(dfl)
    c = d * prim_norm
    p = a[:, None] + a[None, :]
    ovlp = dfl * np.pi ** 1.5 / ((2.0 * p) ** l * p ** 1.5)
    return c / np.sqrt(c @ ovlp @ c)


def _ro(a: np.ndarray) -> np.ndarray:
    a.flags.writeable = False
    return a


@dataclass(frozen=True, eq=False)
class Shell:
    atom: int
    l: int
    center: np.ndarray
    exponents: np.ndarray
    coefficients: np.ndarray
    ao_offset: int

    @property
    def ncart(self) -> int:
        return ncart(self.l)


class BasisSet:
    """Cartesian Gaussian basis of one molecule (see module docstring)."""

    def __init__(self, name: str, symbols: Sequence[str], positions: np.ndarray,
                 shell_atom, shell_l, shell_prim, prim_exp, prim_coef) -> None:
        self.name = name
        self.symbols = [normalize_symbol(s) for s in symbols]
        self.natm = len(self.symbols)
        self.atomic_numbers = _ro(np.array([ATOMIC_NUMBERS[s] for s in self.symbols], dtype=np.int64))
        self.shell_atom = _ro(np.asarray(shell_atom, dtype=np.int64).copy())
        self.shell_l = _ro(np.asarray(shell_l, dtype=np.int64).copy())
        self.shell_prim = _ro(np.asarray(shell_prim, dtype=np.int64).copy())
        self.prim_exp = _ro(np.asarray(prim_exp, dtype=float).copy())
        self.prim_coef = _ro(np.asarray(prim_coef, dtype=float).copy())
        self.nshell = int(self.shell_l.shape[0])
        self.nprim = int(self.prim_exp.shape[0])
        nc = (self.shell_l + 1) * (self.shell_l + 2) // 2
        self.shell_ao = _ro(np.concatenate([[0], np.cumsum(nc)]).astype(np.int64))
        self.nao = int(self.shell_ao[-1])
        self.lmax = int(self.shell_l.max()) if self.nshell else 0
        self.ao_shell = _ro(np.repeat(np.arange(self.nshell, dtype=np.int64), nc))
        self.ao_atom = _ro(self.shell_atom[self.ao_shell].copy())
        self.ao_l = _ro(self.shell_l[self.ao_shell].copy())
        comps = [c for l in self.shell_l for c in cartesian_components(int(l))]
        self.ao_lxyz = _ro(np.array(comps, dtype=np.int64).reshape(-1, 3))
        rng = np.zeros((self.natm, 2), dtype=np.int64)
        for a in range(self.natm):
            idx = np.nonzero(self.ao_atom == a)[0]
            if len(idx):
                rng[a] = (idx[0], idx[-1] + 1)
        self.atom_ao = _ro(rng)
        self._make_blocks()
        self._set_positions(positions)

    def _make_blocks(self) -> None:
        """Group shells into contraction blocks: same atom, same l, nested exponent sets."""
        starts = []
        union: set[float] = set()
        for s in range(self.nshell):
            k0, k1 = self.shell_prim[s], self.shell_prim[s + 1]
            exps = set(self.prim_exp[k0:k1].tolist())
            joins = (s > 0 and self.shell_atom[s] == self.shell_atom[s - 1]
                     and self.shell_l[s] == self.shell_l[s - 1]
                     and (exps <= union or union <= exps))
            if joins:
                union |= exps
            else:
                starts.append(s)
                union = set(exps)
        bs = np.array(starts + [self.nshell], dtype=np.int64)
        self.nblock = len(starts)
        self.blk_shell = _ro(bs)
        self.blk_atom = _ro(self.shell_atom[bs[:-1]].copy())
        self.blk_l = _ro(self.shell_l[bs[:-1]].copy())
        self.blk_ncon = _ro(np.diff(bs))
        self.blk_ao = _ro(self.shell_ao[bs[:-1]].copy())
        self.shell_block = _ro(np.repeat(np.arange(self.nblock, dtype=np.int64), np.diff(bs)))

    def _set_positions(self, positions) -> None:
        x = np.array(positions, dtype=float).reshape(-1, 3)
        if x.shape[0] != self.natm:
            raise ValueError(f"expected {self.natm} positions, got {x.shape[0]}")
        if not np.all(np.isfinite(x)):
            raise ValueError("atomic positions must be finite (got NaN or inf)")
        self.positions = _ro(x)

    def with_positions(self, positions) -> "BasisSet":
        """The same basis on moved atoms; every array except ``positions`` is shared."""
        new = copy.copy(self)
        new._set_positions(positions)
        return new

    @property
    def nuclear_charges(self) -> np.ndarray:
        return self.atomic_numbers.astype(float)

    @property
    def shell_centers(self) -> np.ndarray:
        return self.positions[self.shell_atom]

    def shells(self) -> list[Shell]:
        return [Shell(int(self.shell_atom[s]), int(self.shell_l[s]),
                      self.positions[self.shell_atom[s]].copy(),
                      self.prim_exp[self.shell_prim[s]:self.shell_prim[s + 1]].copy(),
                      self.prim_coef[self.shell_prim[s]:self.shell_prim[s + 1]].copy(),
                      int(self.shell_ao[s])) for s in range(self.nshell)]

    def ao_labels(self) -> list[str]:
        out = []
        for mu in range(self.nao):
            a = int(self.ao_atom[mu])
            lx, ly, lz = (int(v) for v in self.ao_lxyz[mu])
            out.append(f"{a} {self.symbols[a]} {_L_LABELS[lx + ly + lz]}{'x' * lx + 'y' * ly + 'z' * lz}")
        return out

    def __repr__(self) -> str:
        return (f"BasisSet({self.name!r}, natm={self.natm}, nshell={self.nshell}, "
                f"nao={self.nao}, nprim={self.nprim}, nblock={self.nblock})")


def build_basis(symbols: Sequence[str], positions, basis: str) -> BasisSet:
    """Basis ``basis`` (case-insensitive name) for ``symbols`` at ``positions`` (bohr)."""
    name = normalize_basis_name(basis)
    shell_atom, shell_l, shell_prim, prim_exp, prim_coef = [], [], [0], [], []
    for a, sym in enumerate(symbols):
        for l, exps, coefs in sorted(get_basis_shells(name, sym), key=lambda sh: sh[0]):
            for col in range(coefs.shape[1]):
                keep = coefs[:, col] != 0.0
                c = normalized_contraction(l, exps[keep], coefs[keep, col])
                shell_atom.append(a)
                shell_l.append(l)
                prim_exp.extend(exps[keep])
                prim_coef.extend(c)
                shell_prim.append(len(prim_exp))
    return BasisSet(name, symbols, positions, np.array(shell_atom, dtype=np.int64),
                    np.array(shell_l, dtype=np.int64), np.array(shell_prim, dtype=np.int64),
                    np.array(prim_exp, dtype=float), np.array(prim_coef, dtype=float))
