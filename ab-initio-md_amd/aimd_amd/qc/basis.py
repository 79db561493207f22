"""
Basis packing for the device: element templates concatenated into the flat
shell / primitive / AO arrays the gfx950 engine stages with
``aimdhf_plan_create``, plus the contraction-block index its Schwarz
screening works on.

Packing works per element, not per atom. The first time a (basis, element)
combination is seen, its table rows are turned into a template: shells sorted
by l, a general contraction split into one segmented shell per non-zero
column, coefficients normalised, and the block structure of the element
marked. A molecule is then the concatenation of its atoms' templates with
shell, primitive and AO offsets added, which is a handful of numpy
concatenations even for the (H2O)32 clusters (96 atoms).

Output contract, kept identical to the reference (``aimd/qc/basis.py:1-68``)
because AO density matrices cross the backend boundary:

* AO mu = f(lx,ly,lz) x^lx y^ly z^lz sum_k c_k exp(-a_k r^2) centred on its
  atom, all Cartesian components kept (6 d), each with unit self-overlap;
* c_k normalises the x^l component; component (lx,ly,lz) carries
  f = sqrt((2l-1)!! / ((2lx-1)!! (2ly-1)!! (2lz-1)!!)) on top (applied by the
  integral kernels, not stored);
* order: atoms as given, shells of an atom by l (stable), general contraction
  columns in table order, components xx, xy, xz, yy, yz, zz (PySCF order).

Contraction blocks: consecutive shells on one atom with equal l whose
exponent sets nest (one contains the running union of the other) share a
block. The reference evaluates two-electron integrals per block; this engine
evaluates them per segmented shell but screens with the block-pair Schwarz
bound, so it keeps the same set of block quartets. ``blk_prim``, ``blk_exp``
and ``blk_coef`` describe each block's distinct exponents and its
(exponent x contraction) coefficient matrix for callers that read them.
"""

from __future__ import annotations

import copy
import functools
import math
from dataclasses import dataclass
from typing import Iterator, Sequence

import numpy as np

from aimd_amd.elements import ATOMIC_NUMBERS, normalize_symbol
from aimd_amd.qc.basis_data import get_basis_shells, normalize_basis_name

LMAX_DEVICE = 4     # highest shell l the engine takes (l <= 2 class kernels, l = 3, 4 the runtime-l kernels)
_SHELL_LETTERS = "spdfghi"


def ncart(l: int) -> int:
    return (l + 1) * (l + 2) // 2


def cartesian_components(l: int) -> list[tuple[int, int, int]]:
    """(lx, ly, lz) of a shell in PySCF order (lx descending, then ly descending)."""
    out = []
    for lx in range(l, -1, -1):
        out.extend((lx, ly, l - lx - ly) for ly in range(l - lx, -1, -1))
    return out


def _odd_factorial(l: int) -> int:
    """(2l - 1)!! for l >= 0 (1 for l = 0)."""
    return math.prod(range(1, 2 * l, 2))


def cartesian_factor(lx: int, ly: int, lz: int) -> float:
    """Scale of component (lx, ly, lz) relative to the normalised x^l component."""
    return math.sqrt(_odd_factorial(lx + ly + lz)
                     / (_odd_factorial(lx) * _odd_factorial(ly) * _odd_factorial(lz)))


def normalized_contraction(l: int, exponents, coefficients) -> np.ndarray:
    """
    Coefficients of x^l sum_k c_k exp(-a_k r^2) with unit norm, from
    coefficients over normalised primitives. The normalised primitives'
    overlap matrix has the closed form (2 sqrt(a_i a_j) / (a_i + a_j))^(l + 3/2),
    and a primitive's norm is (2a/pi)^(3/4) (4a)^(l/2) / sqrt((2l-1)!!).
    """
    a = np.asarray(exponents, dtype=float)
    d = np.asarray(coefficients, dtype=float)
    gm = np.sqrt(np.outer(a, a))
    s = (2.0 * gm / (a[:, None] + a[None, :])) ** (l + 1.5)
    scale = 1.0 / math.sqrt(float(d @ s @ d))
    norm = (2.0 * a / math.pi) ** 0.75 * (4.0 * a) ** (0.5 * l) / math.sqrt(_odd_factorial(l))
    return d * norm * scale


def _frozen(a) -> np.ndarray:
    a = np.ascontiguousarray(a)
    a.setflags(write=False)
    return a


@dataclass(frozen=True)
class _Template:
    """Packed shells of one element in one basis (positions-free)."""
    shell_l: np.ndarray       # (nshell,)
    shell_nprim: np.ndarray   # (nshell,)
    exps: np.ndarray          # (nprim,)
    coefs: np.ndarray         # (nprim,), normalised


@functools.lru_cache(maxsize=None)
def _template(basis: str, symbol: str) -> _Template:
    return _pack_rows(get_basis_shells(basis, symbol))


def _pack_rows(shells) -> _Template:
    """Template from (l, exponents (nprim,), coefficients (nprim, ncontr)) shells."""
    shells = sorted(shells, key=lambda sh: sh[0])
    ls, nps, es, cs = [], [], [], []
    for l, exps, table in shells:
        for col in table.T:
            used = col != 0.0            # zero-padded general-contraction columns
            ls.append(l)
            nps.append(int(used.sum()))
            es.append(exps[used])
            cs.append(normalized_contraction(l, exps[used], col[used]))
    return _Template(_frozen(np.array(ls, np.int64)), _frozen(np.array(nps, np.int64)),
                     _frozen(np.concatenate(es)), _frozen(np.concatenate(cs)))


def _block_starts(shell_atom, shell_l, shell_prim, prim_exp) -> np.ndarray:
    """Boolean per shell: True where a new contraction block begins."""
    ns = len(shell_l)
    start = np.ones(ns, bool)
    union: set[float] = set()
    for s in range(ns):
        mine = set(prim_exp[shell_prim[s]:shell_prim[s + 1]].tolist())
        if s and shell_atom[s] == shell_atom[s - 1] and shell_l[s] == shell_l[s - 1] and (
                mine.issubset(union) or union.issubset(mine)):
            start[s] = False
            union.update(mine)
        else:
            union = mine
    return start


@dataclass(frozen=True, eq=False)
class Shell:
    """Read-only view of one segmented shell."""
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
    """
    Flat Cartesian basis of one molecule (module docstring). Built from packed
    arrays: ``shell_prim`` has nshell + 1 offsets into ``prim_exp`` /
    ``prim_coef``; positions in bohr. Every array is read-only, and
    :meth:`with_positions` shares all of them but ``positions``.
    """

    def __init__(self, name: str, symbols: Sequence[str], positions, shell_atom, shell_l, shell_prim,
                 prim_exp, prim_coef) -> None:
        self.name = name
        self.symbols = [normalize_symbol(s) for s in symbols]
        self.natm = len(self.symbols)
        self.atomic_numbers = _frozen(np.fromiter((ATOMIC_NUMBERS[s] for s in self.symbols), np.int64,
                                                  count=self.natm))
        self.shell_atom = _frozen(np.array(shell_atom, dtype=np.int64))
        self.shell_l = _frozen(np.array(shell_l, dtype=np.int64))
        self.shell_prim = _frozen(np.array(shell_prim, dtype=np.int64))
        self.prim_exp = _frozen(np.array(prim_exp, dtype=float))
        self.prim_coef = _frozen(np.array(prim_coef, dtype=float))
        self.nshell = self.shell_l.size
        self.nprim = self.prim_exp.size
        width = (self.shell_l + 1) * (self.shell_l + 2) // 2
        ao = np.zeros(self.nshell + 1, np.int64)
        np.cumsum(width, out=ao[1:])
        self.shell_ao = _frozen(ao)
        self.nao = int(ao[-1])
        self.lmax = int(self.shell_l.max(initial=0))
        self.ao_shell = _frozen(np.repeat(np.arange(self.nshell, dtype=np.int64), width))
        self.ao_atom = _frozen(self.shell_atom[self.ao_shell])
        self.ao_l = _frozen(self.shell_l[self.ao_shell])
        table = {l: np.array(cartesian_components(l), np.int64) for l in set(self.shell_l.tolist())}
        self.ao_lxyz = _frozen(np.concatenate([table[l] for l in self.shell_l.tolist()])
                               if self.nshell else np.zeros((0, 3), np.int64))
        # atoms own contiguous AO ranges (shells are grouped by atom)
        lo = np.searchsorted(self.ao_atom, np.arange(self.natm), side="left")
        hi = np.searchsorted(self.ao_atom, np.arange(self.natm), side="right")
        self.atom_ao = _frozen(np.where((hi > lo)[:, None], np.stack([lo, hi], 1), 0).astype(np.int64))
        self._index_blocks()
        self._place(positions)

    def _index_blocks(self) -> None:
        start = _block_starts(self.shell_atom, self.shell_l, self.shell_prim, self.prim_exp)
        first = np.flatnonzero(start)
        self.nblock = first.size
        self.blk_shell = _frozen(np.append(first, self.nshell).astype(np.int64))
        self.shell_block = _frozen(np.cumsum(start) - 1)
        self.blk_atom = _frozen(self.shell_atom[first])
        self.blk_l = _frozen(self.shell_l[first])
        self.blk_ncon = _frozen(np.diff(self.blk_shell))
        self.blk_ao = _frozen(self.shell_ao[first])
        # distinct exponents of each block (first-appearance order) and the
        # coefficient of each in every contraction of the block
        exps, coefs, counts = [], [], []
        width = int(self.blk_ncon.max(initial=1))
        for b in range(self.nblock):
            s0, s1 = self.blk_shell[b], self.blk_shell[b + 1]
            slot: dict[float, int] = {}
            rows: list[np.ndarray] = []
            for c, s in enumerate(range(s0, s1)):
                for e, v in zip(self.prim_exp[self.shell_prim[s]:self.shell_prim[s + 1]].tolist(),
                                self.prim_coef[self.shell_prim[s]:self.shell_prim[s + 1]].tolist()):
                    if e not in slot:
                        slot[e] = len(rows)
                        rows.append(np.zeros(width))
                    rows[slot[e]][c] += v
            exps.extend(slot)
            coefs.extend(rows)
            counts.append(len(rows))
        self.blk_prim = _frozen(np.concatenate([[0], np.cumsum(counts, dtype=np.int64)]).astype(np.int64))
        self.blk_exp = _frozen(np.array(exps, float))
        self.blk_coef = _frozen(np.array(coefs, float).reshape(-1, width))

    def _place(self, positions) -> None:
        x = np.array(positions, dtype=float)
        if x.size != 3 * self.natm:
            raise ValueError(f"expected {self.natm} positions, got {x.size // 3 if x.ndim else 0}")
        x = x.reshape(self.natm, 3)
        if not np.isfinite(x).all():
            raise ValueError("atomic positions must be finite (got NaN or inf)")
        self.positions = _frozen(x)

    def with_positions(self, positions) -> "BasisSet":
        """This basis on moved atoms (bohr); O(1), every other array is shared."""
        moved = copy.copy(self)
        moved._place(positions)
        return moved

    @property
    def nuclear_charges(self) -> np.ndarray:
        return self.atomic_numbers.astype(float)

    @property
    def shell_centers(self) -> np.ndarray:
        return self.positions[self.shell_atom]

    def shells(self) -> list[Shell]:
        out = []
        for s, (a, l, k0, k1, o) in enumerate(zip(self.shell_atom.tolist(), self.shell_l.tolist(),
                                                   self.shell_prim[:-1].tolist(), self.shell_prim[1:].tolist(),
                                                   self.shell_ao[:-1].tolist())):
            out.append(Shell(a, l, self.positions[a].copy(), self.prim_exp[k0:k1].copy(),
                             self.prim_coef[k0:k1].copy(), o))
        return out

    def ao_labels(self) -> list[str]:
        """'<atom> <symbol> <l letter><axes>', e.g. '0 O dxy'."""
        axes = ["x" * int(i) + "y" * int(j) + "z" * int(k) for i, j, k in self.ao_lxyz]
        return [f"{a} {self.symbols[a]} {_SHELL_LETTERS[len(ax)]}{ax}"
                for a, ax in zip(self.ao_atom.tolist(), axes)]

    def __repr__(self) -> str:
        return (f"BasisSet({self.name!r}, natm={self.natm}, nshell={self.nshell}, "
                f"nao={self.nao}, nprim={self.nprim}, nblock={self.nblock})")


def iter_atom_templates(name: str, symbols: Sequence[str]) -> Iterator[tuple[int, _Template]]:
    for a, sym in enumerate(symbols):
        yield a, _template(name, normalize_symbol(sym))


def build_basis(symbols: Sequence[str], positions, basis: str) -> BasisSet:
    """Basis ``basis`` (case-insensitive name) for ``symbols`` at ``positions`` (bohr)."""
    name = normalize_basis_name(basis)
    return _assemble(name, symbols, positions, list(iter_atom_templates(name, symbols)))


def basis_from_table(symbols: Sequence[str], positions, table: dict, name: str = "custom") -> BasisSet:
    """
    Basis from data in the row layout of ``basis_data.json``,
    {symbol: [[l, [exp, c_1, c_2, ...], ...], ...]} (extra coefficient columns
    form a general contraction; zeros pad), laid out as :func:`build_basis`
    lays out a named basis.
    """
    cache: dict[str, _Template] = {}
    parts = []
    for a, sym in enumerate(symbols):
        key = normalize_symbol(sym)
        if key not in table:
            raise ValueError(f"basis table has no data for element {key!r}")
        if key not in cache:
            rows = [(int(e[0]), np.array(e[1:], float)) for e in table[key]]
            cache[key] = _pack_rows([(l, r[:, 0].copy(), r[:, 1:].copy()) for l, r in rows])
        parts.append((a, cache[key]))
    return _assemble(name, symbols, positions, parts)


def _assemble(name: str, symbols, positions, parts) -> BasisSet:
    if not parts:
        raise ValueError("a basis needs at least one atom")
    shell_atom = np.concatenate([np.full(t.shell_l.size, a, np.int64) for a, t in parts])
    shell_l = np.concatenate([t.shell_l for _, t in parts])
    nprim = np.concatenate([t.shell_nprim for _, t in parts])
    shell_prim = np.zeros(nprim.size + 1, np.int64)
    np.cumsum(nprim, out=shell_prim[1:])
    return BasisSet(name, symbols, positions, shell_atom, shell_l, shell_prim,
                    np.concatenate([t.exps for _, t in parts]), np.concatenate([t.coefs for _, t in parts]))
