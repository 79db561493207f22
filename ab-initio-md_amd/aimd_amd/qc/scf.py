"""
Restricted (RHF) and unrestricted (UHF) Hartree-Fock SCF on the GPU.

Mirrors the reference's solver (aimd/qc/scf.py) -- the same options and
defaults, guesses, DIIS, convergence test and result fields -- with every
matrix held in float64 PyTorch tensors on the HIP device and the two-electron
part computed by the gfx950 engine (aimd_amd.qc.engine), integral-direct:
no ERI tensor is stored, so there is no nao limit of the reference's kind.

Equations (aimd/qc/scf.py:9-23, 80-104): spin densities D_s, g = 2/nspin,
    F_s = h + g sum_s' J[D_s'] - K[D_s],   E_el = 1/2 g sum_s <D_s, h + F_s>,
    e_s = X^T (F_s D_s S - S D_s F_s) X,   converged iff |dE| < conv_energy
                                           and max|e| < conv_commutator.
The returned state is D_k, F[D_k], E[D_k] of the last iteration; the MOs are
the canonical orbitals of F_k.

One host synchronisation per iteration reads the two convergence scalars;
DIIS coefficients come from the (m x m) Gram matrix of the error
differences, formed on the device and solved on the host (m <= diis_space).
"""

from __future__ import annotations

import ctypes
import time
import warnings
from dataclasses import dataclass, field, fields, replace
from typing import Any, Sequence

import numpy as np
import torch

from aimd_amd.elements import normalize_symbol
from aimd_amd import _lib
from aimd_amd.qc import engine as qengine
from aimd_amd.qc.basis import BasisSet, build_basis

LINDEP_THRESHOLD = 1e-7
GWH_K = 1.75
MIN_NUCLEAR_SEPARATION = 1e-5
_GUESSES = ("sad", "core", "gwh")
_ORTHO = ("auto", "symmetric", "canonical")
_DENSITY_SOLVERS = ("purify", "eigh")
_REFERENCES = {"rhf": "rhf", "restricted": "rhf", "uhf": "uhf", "unrestricted": "uhf"}


class SCFConvergenceWarning(RuntimeWarning):
    """Emitted when an SCF stops at ``max_iter`` without converging."""


@dataclass(frozen=True)
class SCFOptions:
    """
    SCF settings; names, defaults and validation as aimd/qc/scf.py:185-247.
    ``max_nao`` and ``blas_threads`` are accepted for compatibility: the
    engine never forms the nao^4 tensor (nothing to cap) and runs no BLAS
    threads. ``jk_mode`` ("auto", "direct", "store") and ``max_store_bytes``
    have no counterpart in the reference: they choose between recomputing the
    screened integrals every J/K build and evaluating them once per geometry
    into device memory (auto: store when it fits in max_store_bytes, 0 = half
    the free device memory). ``incremental_fock`` (default: on in mixed
    precision, off in FP64) builds J and K of every iteration after the first
    from the density difference, J(D) = J(D_prev) + J(D - D_prev), skipping in
    store mode the work units whose largest Schwarz product times max |dD| is
    below ``incremental_threshold`` (default: the Schwarz threshold in FP64; in
    mixed precision a tenth of ``conv_energy`` when that is larger, so the
    skipped contributions stay an order below the SCF's own energy tolerance)
    and, in mixed precision once max |dD| <= ``delta_fp32_switch``, digesting
    the difference in FP32.
    """
    conv_energy: float = 1e-10
    conv_commutator: float = 1e-7
    max_iter: int = 100
    guess: str = "sad"
    diis: bool = True
    diis_space: int = 8
    diis_restart: int = 12
    level_shift: float = 0.0
    damping: float = 0.0
    stabilize_until: float = 1e-3
    orthogonalization: str = "auto"
    lindep_threshold: float = LINDEP_THRESHOLD
    schwarz_threshold: float = qengine.DEFAULT_SCHWARZ
    max_nao: int = 120
    warn_unconverged: bool = True
    blas_threads: int | None = 1
    jk_mode: str = "auto"
    max_store_bytes: int = 0
    density_solver: str = "purify"
    precision: str = "fp64"
    jk_fp32_switch: float = 1e-5
    grad_fp32_switch: float = 1e-5
    delta_fp32_switch: float = 1e-2
    incremental_fock: bool | None = None
    incremental_threshold: float | None = None

    def __post_init__(self) -> None:
        if not (self.conv_energy > 0.0 and self.conv_commutator > 0.0):
            raise ValueError("SCF convergence thresholds must be positive")
        if self.max_iter < 1:
            raise ValueError("max_iter must be >= 1")
        if self.guess not in _GUESSES:
            raise ValueError(f"unknown guess {self.guess!r}; use one of {_GUESSES}")
        if self.diis_space < 2:
            raise ValueError("diis_space must be >= 2")
        if self.diis_restart < 0:
            raise ValueError("diis_restart must be >= 0")
        if self.level_shift < 0.0:
            raise ValueError("level_shift must be >= 0")
        if not 0.0 <= self.damping < 1.0:
            raise ValueError("damping must be in [0, 1)")
        if self.orthogonalization not in _ORTHO:
            raise ValueError(f"unknown orthogonalization {self.orthogonalization!r}; use one of {_ORTHO}")
        if not self.lindep_threshold > 0.0:
            raise ValueError("lindep_threshold must be positive")
        if not self.schwarz_threshold >= 0.0:
            raise ValueError("schwarz_threshold must be >= 0")
        if self.blas_threads is not None and self.blas_threads < 1:
            raise ValueError("blas_threads must be >= 1 or None")
        if self.jk_mode not in qengine.Engine.JK_MODES:
            raise ValueError(f"unknown jk_mode {self.jk_mode!r}; use one of {sorted(qengine.Engine.JK_MODES)}")
        if self.max_store_bytes < 0:
            raise ValueError("max_store_bytes must be >= 0")
        if self.precision not in qengine.Engine.PRECISIONS:
            raise ValueError(f"unknown precision {self.precision!r}; use one of {sorted(qengine.Engine.PRECISIONS)}")
        if not (self.jk_fp32_switch >= 0.0 and self.grad_fp32_switch >= 0.0 and self.delta_fp32_switch >= 0.0):
            raise ValueError("fp32 switch thresholds must be >= 0")
        if self.incremental_threshold is not None and not (self.incremental_threshold >= 0.0
                                                            and np.isfinite(self.incremental_threshold)):
            raise ValueError("incremental_threshold must be finite and >= 0")
        if self.density_solver not in _DENSITY_SOLVERS:
            raise ValueError(f"unknown density_solver {self.density_solver!r}; use one of {_DENSITY_SOLVERS}")


def electron_counts(nuclear_charge: int, charge: int, multiplicity: int,
                    reference: str = "uhf") -> tuple[int, int]:
    """(n_alpha, n_beta); ValueError for impossible combinations (aimd/qc/scf.py:250-279)."""
    n = int(nuclear_charge) - int(charge)
    mult = int(multiplicity)
    if n <= 0:
        raise ValueError(f"charge {charge} leaves {n} electrons; need at least one")
    if mult < 1:
        raise ValueError(f"multiplicity must be >= 1, got {mult}")
    if (n + mult - 1) % 2:
        raise ValueError(
            f"multiplicity {mult} is impossible with {n} electrons "
            f"({'even' if n % 2 == 0 else 'odd'} electron counts need "
            f"{'odd' if n % 2 == 0 else 'even'} multiplicities)")
    if mult - 1 > n:
        raise ValueError(f"multiplicity {mult} needs at least {mult - 1} electrons, have {n}")
    na, nb = (n + mult - 1) // 2, (n - mult + 1) // 2
    ref = str(reference).strip().lower()
    if ref not in _REFERENCES:
        raise ValueError(f"unknown reference {reference!r}; use 'rhf' or 'uhf'")
    if _REFERENCES[ref] == "rhf" and na != nb:
        raise ValueError(f"RHF needs a closed-shell singlet; {n} electrons with multiplicity "
                         f"{mult} is open-shell (use UHF)")
    return na, nb


def check_nuclear_separation(positions: np.ndarray, atomic_numbers: np.ndarray) -> None:
    """ValueError if two nuclei are closer than 1e-5 bohr (aimd/qc/scf.py:666-684)."""
    x = np.asarray(positions, dtype=float).reshape(-1, 3)
    z = np.asarray(atomic_numbers)
    if len(x) < 2 or not np.all(np.isfinite(x)):
        return
    r = np.linalg.norm(x[:, None, :] - x[None, :, :], axis=-1)
    i, j = np.triu_indices(len(x), k=1)
    bad = (r[i, j] < MIN_NUCLEAR_SEPARATION) & (z[i] > 0) & (z[j] > 0)
    if np.any(bad):
        k = int(np.argmax(bad))
        raise ValueError(
            f"atoms {i[k]} and {j[k]} are {r[i[k], j[k]]:.3g} bohr apart (coincident "
            f"nuclei, < {MIN_NUCLEAR_SEPARATION:g} bohr): the nuclear repulsion is infinite")


def purified_projector(F: torch.Tensor, nocc: int, tol: float = 1e-11, max_iter: int = 120,
                       check_every: int = 8) -> torch.Tensor | None:
    """
    Projector onto the eigenvectors of the nocc lowest eigenvalues of symmetric F
    by trace-correcting purification (Niklasson, Phys. Rev. B 66, 155115 (2002)):
    X0 = (e_max - F) / (e_max - e_min) with Gershgorin bounds, then X <- X^2 while
    tr X > nocc, else X <- 2X - X^2. Branches are chosen on the device; the
    idempotency error tr(X - X^2) is read back every check_every steps. Once it
    is below tol, two more steps square the remaining error away. Returns None if
    it has not converged after max_iter steps (no gap at the Fermi level).
    """
    n = F.shape[0]
    eye = torch.eye(n, dtype=F.dtype, device=F.device)
    if nocc <= 0:
        return torch.zeros_like(F)
    if nocc >= n:
        return eye
    a = F.abs()
    d = torch.diagonal(F)
    r = a.sum(dim=1) - torch.diagonal(a)
    emin, emax = (d - r).min(), (d + r).max()
    X = (emax * eye - F) / (emax - emin)
    it = 0
    while it < max_iter:
        for _ in range(check_every):
            X2 = X @ X
            X = torch.where(torch.trace(X) > nocc, X2, 2.0 * X - X2)
        it += check_every
        err = float((torch.trace(X) - torch.sum(X * X)).abs())
        if err < tol:
            for _ in range(2):
                X2 = X @ X
                X = torch.where(torch.trace(X) > nocc, X2, 2.0 * X - X2)
            return 0.5 * (X + X.T)
    return None


def inverse_sqrt_newton_schulz(S: torch.Tensor, max_iter: int = 60, tol: float = 1e-12,
                               check_every: int = 6) -> tuple[torch.Tensor, float] | None:
    """
    S^{-1/2} of a symmetric positive definite S by the coupled Newton-Schulz
    iteration (Higham, Functions of Matrices, eq. 6.35): Y0 = S/c, Z0 = 1,
    T = (3 - Z Y)/2, Y <- Y T, Z <- T Z, with c a Gershgorin bound on the largest
    eigenvalue; Z -> (S/c)^{-1/2}. Returns (S^{-1/2}, a lower bound of the smallest
    eigenvalue of S: 1 / max_i sum_j |(S^{-1})_ij|, since the largest eigenvalue of
    S^{-1} = X X is at most its largest absolute row sum), or None if |X S X - 1|
    does not fall below 1e-10 within max_iter steps.
    """
    n = S.shape[0]
    eye = torch.eye(n, dtype=S.dtype, device=S.device)
    c = S.abs().sum(dim=1).max()
    Y = S / c
    Z = eye.clone()
    for it in range(1, max_iter + 1):
        T = 0.5 * (3.0 * eye - Z @ Y)
        Y = Y @ T
        Z = T @ Z
        if it % check_every == 0 and float((Z @ Y - eye).abs().max()) < tol:
            break
    X = Z / torch.sqrt(c)
    X = 0.5 * (X + X.T)
    if not float((X @ S @ X - eye).abs().max()) < 1e-10:
        return None
    lam = float((X @ X).abs().sum(dim=1).max())
    return X, 1.0 / lam


class TC2Graph:
    """
    Purification steps of an n x n matrix with nocc electrons captured once as HIP
    graphs, STEPS steps (``replay``) and TAIL_STEPS steps (``tail``) on the same
    buffers: each step is a dense FP64 product X2 = X X (BLAS) and the engine's
    fused update (aimdhf_tc2_step), ping-ponging between two static buffers so the
    result of an even number of steps lands back in ``X``. The update kernel needs
    no reduction pass and clears no memory, so replays are exact repeats. The
    replayed steps hold an iterate whose idempotency error is below ``tol``
    (steps past convergence would amplify its rounding errors until they
    overflow); the tail steps do not, they square the remaining error away.
    ``needed`` remembers how many replays the last purification took; the next one
    (the Fock matrix of the next SCF iteration) checks convergence from one replay
    earlier on, so the count follows the matrices down as well as up.
    """
    STEPS = 8
    TAIL_STEPS = 2

    def __init__(self, n: int, nocc: int, device: torch.device, tol: float = 1e-11) -> None:
        self.n, self.nocc, self.tol = n, float(nocc), float(tol)
        self.needed = 1
        self.X = torch.zeros((n, n), dtype=torch.float64, device=device)
        self.B = torch.empty_like(self.X)
        self.X2 = torch.empty_like(self.X)
        self.err = torch.zeros(1, dtype=torch.float64, device=device)
        side = torch.cuda.Stream(device=device)
        side.wait_stream(torch.cuda.current_stream(device))
        with torch.cuda.stream(side):
            self._steps(self.STEPS, self.tol)
        torch.cuda.current_stream(device).wait_stream(side)
        self.graph = torch.cuda.CUDAGraph()
        with torch.cuda.graph(self.graph):
            self._steps(self.STEPS, self.tol)
        self.tail_graph = torch.cuda.CUDAGraph()
        with torch.cuda.graph(self.tail_graph):
            self._steps(self.TAIL_STEPS, 0.0)

    def _steps(self, nsteps: int, freeze: float) -> None:
        L = _lib.lib()
        st = ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)
        src, dst = self.X, self.B
        for _ in range(nsteps):
            torch.matmul(src, src, out=self.X2)
            _lib.check(L.aimdhf_tc2_step(self.n, ctypes.c_void_p(src.data_ptr()), ctypes.c_void_p(self.X2.data_ptr()),
                                         ctypes.c_void_p(dst.data_ptr()), self.nocc, freeze, st))
            src, dst = dst, src

    def replay(self) -> torch.Tensor:
        self.graph.replay()
        return self.X

    def tail(self) -> torch.Tensor:
        self.tail_graph.replay()
        return self.X

    def start(self, F: torch.Tensor) -> None:
        """X = (e_max - F) / (e_max - e_min), Gershgorin bounds of F (one kernel)."""
        F = F.contiguous()
        _lib.check(_lib.lib().aimdhf_tc2_init(self.n, ctypes.c_void_p(F.data_ptr()), ctypes.c_void_p(self.X.data_ptr()),
                                              ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)))

    def error(self) -> float:
        """|tr X - sum X_ij^2| of the current iterate (one kernel and one read)."""
        _lib.check(_lib.lib().aimdhf_tc2_error(self.n, ctypes.c_void_p(self.X.data_ptr()),
                                               ctypes.c_void_p(self.err.data_ptr()),
                                               ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)))
        return float(self.err)


_TC2_GRAPHS: dict[tuple, TC2Graph] = {}


def purified_projector_device(F: torch.Tensor, nocc: int, tol: float = 1e-11,
                              max_replays: int = 15) -> torch.Tensor | None:
    """purified_projector on the GPU with the captured TC2Graph (same steps, same branch
    rule): the idempotency error is read back after the number of replays the previous
    call needed minus one, then after every replay, and the TAIL_STEPS tail graph squares
    the remaining error away. Replayed steps hold a converged iterate unchanged, so
    starting the checks late costs time, not accuracy; starting one replay early lets
    the count drop again after a hard matrix."""
    n = F.shape[0]
    if nocc <= 0:
        return torch.zeros_like(F)
    if nocc >= n:
        return torch.eye(n, dtype=F.dtype, device=F.device)
    key = (n, int(nocc), F.device, float(tol))
    g = _TC2_GRAPHS.get(key)
    if g is None:
        g = _TC2_GRAPHS[key] = TC2Graph(n, nocc, F.device, tol)
    g.start(F)
    first = max(1, min(g.needed - 1, max_replays))
    for _ in range(first - 1):
        g.replay()
    for k in range(first, max_replays + 1):
        g.replay()
        if g.error() < tol:
            g.needed = k
            X = g.tail()
            return 0.5 * (X + X.T)
    g.needed = 1
    return None


def orthogonalizer(S: torch.Tensor, method: str = "auto",
                   threshold: float = LINDEP_THRESHOLD) -> tuple[torch.Tensor, int, float]:
    """
    X with X^T S X = 1, removed combinations, smallest eigenvalue (aimd/qc/scf.py:284-303).
    The symmetric S^{-1/2} is formed by Newton-Schulz iteration (dense FP64 products)
    when the smallest eigenvalue is clearly above the threshold (a lower bound of it
    > 1e3 x threshold; that bound is then the third value returned); otherwise, and
    for canonical orthogonalisation, from eigh as the reference does, with its
    linear-dependence handling.
    """
    if method not in _ORTHO:
        raise ValueError(f"unknown orthogonalization {method!r}; use one of {_ORTHO}")
    if method in ("auto", "symmetric"):
        ns = inverse_sqrt_newton_schulz(S)
        if ns is not None and ns[1] > 1e3 * threshold:
            return ns[0], 0, ns[1]
    s, U = torch.linalg.eigh(S)
    s_h = s.cpu().numpy()
    smin = float(s_h[0])
    keep = s_h >= threshold
    if method == "symmetric" or (method == "auto" and keep.all()):
        if not keep.all():
            raise ValueError(
                f"overlap eigenvalue {smin:.3e} < lindep_threshold {threshold:.1e}: "
                "the basis is nearly linearly dependent; use canonical orthogonalization")
        return (U * s.rsqrt()) @ U.T, 0, smin
    if not keep.any():
        raise ValueError("every overlap eigenvalue is below lindep_threshold")
    idx = torch.as_tensor(np.nonzero(keep)[0], device=S.device)
    return U[:, idx] * s[idx].rsqrt(), int((~keep).sum()), smin


def nuclear_repulsion_energy(charges: np.ndarray, positions: np.ndarray) -> float:
    q = np.asarray(charges, dtype=float).reshape(-1)
    x = np.asarray(positions, dtype=float).reshape(-1, 3)
    i, j = np.triu_indices(len(q), k=1)
    return float(np.sum(q[i] * q[j] / np.linalg.norm(x[i] - x[j], axis=1)))


def nuclear_repulsion_gradient(charges: np.ndarray, positions: np.ndarray) -> np.ndarray:
    q = np.asarray(charges, dtype=float).reshape(-1)
    x = np.asarray(positions, dtype=float).reshape(-1, 3)
    i, j = np.triu_indices(len(q), k=1)
    d = x[i] - x[j]
    r = np.linalg.norm(d, axis=1)
    pair = -(q[i] * q[j] / r ** 3)[:, None] * d
    g = np.zeros_like(x)
    np.add.at(g, i, pair)
    np.add.at(g, j, -pair)
    return g


def nuclear_dipole(charges, positions, origin=None) -> np.ndarray:
    q = np.asarray(charges, dtype=float).reshape(-1)
    x = np.asarray(positions, dtype=float).reshape(-1, 3)
    o = np.zeros(3) if origin is None else np.asarray(origin, dtype=float).reshape(3)
    return q @ (x - o)


# ================================================================ integrals
@dataclass(eq=False)
class SCFIntegrals:
    """One-electron matrices and the engine state of one geometry (device tensors)."""
    basis: BasisSet
    engine: qengine.Engine
    S: torch.Tensor
    H: torch.Tensor
    M: torch.Tensor          # dipole integrals (3, nao, nao) about the origin
    X: torch.Tensor
    n_removed: int
    min_overlap_eigenvalue: float
    nuclear_repulsion: float

    @property
    def nao(self) -> int:
        return self.basis.nao

    @property
    def nmo(self) -> int:
        return int(self.X.shape[1])


def build_integrals(basis: BasisSet, engine: qengine.Engine, options: SCFOptions) -> SCFIntegrals:
    """Geometry set-up, S, h = T + V, dipole integrals and X (aimd/qc/scf.py:329-336)."""
    engine.set_geometry(basis.positions, options.schwarz_threshold)
    n = basis.nao
    dev = engine.device
    S = torch.empty((n, n), dtype=torch.float64, device=dev)
    T = torch.empty_like(S)
    V = torch.empty_like(S)
    M = torch.empty((3, n, n), dtype=torch.float64, device=dev)
    engine.one_electron(S, T, V, M)
    X, n_removed, smin = orthogonalizer(S, options.orthogonalization, options.lindep_threshold)
    enuc = nuclear_repulsion_energy(basis.nuclear_charges, basis.positions)
    return SCFIntegrals(basis, engine, S, T + V, M, X, n_removed, smin, enuc)


# ================================================================ guesses
_MADELUNG = sorted(((n, l) for n in range(1, 8) for l in range(n)), key=lambda t: (t[0] + t[1], t[0]))


def ground_state_l_counts(Z: int) -> list[int]:
    counts = [0, 0, 0, 0]
    left = int(Z)
    for _n, l in _MADELUNG:
        if left <= 0:
            break
        k = min(left, 2 * (2 * l + 1))
        counts[l] += k
        left -= k
    return counts


def _degenerate_groups(eps: np.ndarray, tol: float) -> list[tuple[int, int]]:
    groups, i = [], 0
    while i < len(eps):
        j = i + 1
        while j < len(eps) and eps[j] - eps[j - 1] < tol:
            j += 1
        groups.append((i, j))
        i = j
    return groups


def _atomic_occupations(eps: np.ndarray, l_counts: Sequence[int], tol: float = 1e-6) -> np.ndarray:
    """Spin-summed shell occupations of a spherical atom (aimd/qc/scf.py:394-424)."""
    groups = _degenerate_groups(eps, tol)
    occ = np.zeros(len(eps))
    left = list(l_counts) + [0] * (5 - len(l_counts))
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
    n = float(sum(l_counts))
    for i, j in groups:
        k = min(n, 2.0 * (j - i))
        occ[i:j] = k / (j - i)
        n -= k
    return occ


def atom_basis(basis: BasisSet, atom: int) -> BasisSet:
    """Shells of one atom as a one-atom basis at the origin (aimd/qc/scf.py:492-501)."""
    sh = np.nonzero(basis.shell_atom == atom)[0]
    k0 = [basis.shell_prim[s] for s in sh]
    k1 = [basis.shell_prim[s + 1] for s in sh]
    prim = np.concatenate([np.arange(a, b) for a, b in zip(k0, k1)]) if len(sh) else np.zeros(0, int)
    shell_prim = np.concatenate([[0], np.cumsum([b - a for a, b in zip(k0, k1)])])
    return BasisSet(basis.name, [basis.symbols[atom]], np.zeros((1, 3)),
                    np.zeros(len(sh), dtype=np.int64), basis.shell_l[sh], shell_prim,
                    basis.prim_exp[prim], basis.prim_coef[prim])


_ATOMIC_CACHE: dict[tuple, tuple[np.ndarray, float, bool]] = {}


def atomic_hf(basis_atom: BasisSet, max_iter: int = 200) -> tuple[np.ndarray, float, bool]:
    """
    Spherically averaged fractional-occupation RHF of a neutral atom in its
    molecular basis (aimd/qc/scf.py:430-484), on the device; cached per
    element and shell data.
    """
    if basis_atom.natm != 1:
        raise ValueError("atomic_hf needs a one-atom basis")
    key = (int(basis_atom.atomic_numbers[0]), basis_atom.shell_l.tobytes(), basis_atom.shell_prim.tobytes(),
           basis_atom.prim_exp.tobytes(), basis_atom.prim_coef.tobytes())
    hit = _ATOMIC_CACHE.get(key)
    if hit is not None:
        return hit[0].copy(), hit[1], hit[2]
    Z = int(basis_atom.atomic_numbers[0])
    lc = ground_state_l_counts(Z)
    eng = qengine.Engine(basis_atom)
    try:
        I = build_integrals(basis_atom, eng, SCFOptions(orthogonalization="canonical"))
        S, H = I.S, I.H
        X, _, _ = orthogonalizer(S, "canonical", LINDEP_THRESHOLD)
        if Z == 1:
            eps, Cp = torch.linalg.eigh(X.T @ H @ X)
            c = X @ Cp[:, 0]
            D = torch.outer(c, c).cpu().numpy()
            res = (D, float(eps[0]), True)
            _ATOMIC_CACHE[key] = res
            return D.copy(), res[1], True

        def density(Fo: torch.Tensor) -> torch.Tensor:
            eps, Cp = torch.linalg.eigh(Fo)
            C = X @ Cp
            occ = torch.as_tensor(_atomic_occupations(eps.cpu().numpy(), lc), device=C.device)
            return (C * occ) @ C.T

        D = density(X.T @ H @ X)
        diis = DIIS(8)
        e_old = np.inf
        converged = False
        e = 0.0
        for _ in range(max_iter):
            J, K = eng.jk(D)
            F = H + J - 0.5 * K
            F = 0.5 * (F + F.T)
            e = 0.5 * float(torch.vdot(D.reshape(-1), (H + F).reshape(-1)))
            err = X.T @ (F @ D @ S - S @ D @ F) @ X
            if abs(e - e_old) < 1e-11 and float(err.abs().max()) < 1e-8:
                converged = True
                break
            e_old = e
            diis.push((X.T @ F @ X)[None], err[None])
            D = density(diis.extrapolate()[0])
        Dn = (0.5 * (D + D.T)).cpu().numpy()
    finally:
        eng.close()
    _ATOMIC_CACHE[key] = (Dn, e, converged)
    return Dn.copy(), e, converged


def sad_density(basis: BasisSet) -> np.ndarray:
    """Superposition of atomic densities (total density, (nao, nao)) (aimd/qc/scf.py:504-511)."""
    P = np.zeros((basis.nao, basis.nao))
    for a in range(basis.natm):
        i, j = basis.atom_ao[a]
        if j > i:
            P[i:j, i:j] = atomic_hf(atom_basis(basis, a))[0]
    return P


# =================================================================== DIIS
class DIIS:
    """
    Pulay DIIS anchored on the newest vector (aimd/qc/scf.py:516-568):
    minimise |e_m + sum_i c_i (e_i - e_m)|; the minimum-norm solution is taken
    from the eigen-decomposition of the small Gram matrix of the differences
    (singular values below eps max(M, N) of the largest are dropped, numpy's
    lstsq rule), and the oldest vectors are dropped while |c| > 1e4.

    The vectors live in two ring buffers on the device of the first push; each
    push copies its pair in and adds one row of the raw Gram matrix (one GEMV),
    from which the Gram matrix of the differences follows without touching the
    vectors again; the extrapolation is one GEMV over the ring.
    """

    def __init__(self, space: int = 8, max_coefficient: float = 1e4) -> None:
        self.space = int(space)
        self.max_coefficient = float(max_coefficient)
        self._F: torch.Tensor | None = None     # (space, *fock.shape)
        self._E: torch.Tensor | None = None     # (space, *error.shape)
        self._G: torch.Tensor | None = None     # (space, space) raw Gram <e_i, e_j> of the ring slots
        self._head = 0                          # slot of the oldest vector
        self._m = 0                             # vectors held (slots head .. head+m-1 mod space)

    def __len__(self) -> int:
        return self._m

    def _slot(self, k: int) -> int:
        return (self._head + k) % self.space

    @property
    def focks(self) -> list[torch.Tensor]:
        return [self._F[self._slot(k)] for k in range(self._m)]

    @property
    def errors(self) -> list[torch.Tensor]:
        return [self._E[self._slot(k)] for k in range(self._m)]

    def reset(self) -> None:
        self._m = 0
        if self._F is not None:
            self._F.zero_()
            self._E.zero_()

    def push(self, fock: torch.Tensor, error: torch.Tensor) -> None:
        if self._F is None or self._F.shape[1:] != fock.shape or self._F.device != fock.device:
            self._F = torch.zeros((self.space,) + tuple(fock.shape), dtype=fock.dtype, device=fock.device)
            self._E = torch.zeros((self.space,) + tuple(error.shape), dtype=error.dtype, device=error.device)
            self._G = torch.zeros((self.space, self.space), dtype=error.dtype, device=error.device)
            self._head, self._m = 0, 0
        if self._m == self.space:
            self._head = (self._head + 1) % self.space
            self._m -= 1
        slot = self._slot(self._m)
        self._F[slot].copy_(fock)
        self._E[slot].copy_(error)
        g = self._E.view(self.space, -1) @ error.reshape(-1)
        self._G[slot] = g
        self._G[:, slot] = g
        self._m += 1

    def keep_last(self) -> None:
        """Restart from the newest vector alone."""
        self._head = self._slot(self._m - 1)
        self._m = 1

    def gram(self) -> torch.Tensor | None:
        """Device (m, m-1) array: rows 0..m-2 the Gram matrix of the differences
        e_i - e_m, row m-1 the right-hand side <e_i - e_m, -e_m>; None for m = 1.
        Computed before the host needs it, so it travels with other scalars."""
        m = self._m
        if m == 1:
            return None
        Gs = torch.roll(self._G, shifts=(-self._head, -self._head), dims=(0, 1))[:m, :m]
        a = m - 1
        Gd = Gs[:a, :a] - Gs[:a, a:] - Gs[a:, :a] + Gs[a, a]
        r = Gs[a, a] - Gs[a, :a]
        return torch.cat([Gd, r[None]])

    def gram_from(self, G_raw: np.ndarray) -> np.ndarray | None:
        """gram() from a host copy of the raw (space, space) slot Gram matrix, which
        travels with the other per-iteration scalars: the same differences, formed
        on the host."""
        m = self._m
        if m == 1:
            return None
        Gs = np.roll(G_raw, shift=(-self._head, -self._head), axis=(0, 1))[:m, :m]
        a = m - 1
        Gd = Gs[:a, :a] - Gs[:a, a:] - Gs[a:, :a] + Gs[a, a]
        r = Gs[a, a] - Gs[a, :a]
        return np.concatenate([Gd, r[None]])

    @property
    def raw_gram(self) -> torch.Tensor:
        """The device (space, space) Gram matrix of the ring slots."""
        return self._G

    def coefficients(self, Gr: np.ndarray | None = None) -> np.ndarray:
        """Coefficients of the stored vectors (oldest first); Gr is the host copy of
        gram() for the current vectors, computed here when not given."""
        m = self._m
        if m == 1:
            return np.ones(1)
        if Gr is None:
            Gr = self.gram().cpu().numpy()
        G, r = Gr[:-1], Gr[-1]
        lam, V = np.linalg.eigh(G)
        sig = np.sqrt(np.maximum(lam, 0.0))
        cut = np.finfo(float).eps * max(m - 1, self._E[0].numel()) * sig.max()
        keep = sig > cut
        x = V[:, keep] @ ((V[:, keep].T @ r) / lam[keep]) if keep.any() else np.zeros(m - 1)
        return np.append(x, 1.0 - x.sum())

    def extrapolate(self, Gr: np.ndarray | None = None) -> torch.Tensor:
        """sum_i c_i F_i; dropping the oldest vector keeps the newest as the anchor,
        so the Gram matrix of the rest is Gr[1:, 1:]."""
        while self._m > 1:
            c = self.coefficients(Gr)
            if np.all(np.isfinite(c)) and np.abs(c).max() <= self.max_coefficient:
                full = np.zeros(self.space)
                for k in range(self._m):
                    full[self._slot(k)] = c[k]
                ct = torch.from_numpy(full)
                if self._F.is_cuda:
                    ct = ct.pin_memory().to(self._F.device, non_blocking=True)
                return torch.mv(self._F.view(self.space, -1).t(), ct).view(self._F.shape[1:])
            self._head = (self._head + 1) % self.space
            self._m -= 1
            if Gr is not None:
                Gr = Gr[1:, 1:]
        return self._F[self._slot(self._m - 1)].clone()


# ================================================================ results
@dataclass(eq=False)
class SCFResult:
    """Converged (or flagged) SCF state; numpy arrays as aimd/qc/scf.py:573-656."""
    reference: str
    converged: bool
    iterations: int
    energy: float
    electronic_energy: float
    nuclear_repulsion: float
    spin_densities: np.ndarray
    fock: np.ndarray
    mulliken_charges: np.ndarray
    dipole: np.ndarray
    s2: float
    n_alpha: int
    n_beta: int
    energy_change: float
    commutator_norm: float
    basis: BasisSet
    n_removed: int = 0
    guess: str = ""
    diis_restarts: int = 0
    history: list[dict[str, float]] = field(default_factory=list)
    timings: dict[str, float] = field(default_factory=dict)
    # device state of the last iteration (for the gradient)
    integrals: SCFIntegrals | None = None
    device_spin_densities: torch.Tensor | None = None   # (nspin, nao, nao)
    device_fock: torch.Tensor | None = None             # (nspin, nao, nao)
    # canonical orbitals (mo_energy, mo_coeff, mo_occ, energy_weighted_density):
    # diagonalised on first access, the MD path never needs them
    orbitals: Any = None

    def _orbitals(self) -> tuple[np.ndarray, np.ndarray, np.ndarray, np.ndarray]:
        if callable(self.orbitals):
            self.orbitals = self.orbitals()
        return self.orbitals

    @property
    def mo_energy(self) -> np.ndarray:
        return self._orbitals()[0]

    @property
    def mo_coeff(self) -> np.ndarray:
        return self._orbitals()[1]

    @property
    def mo_occ(self) -> np.ndarray:
        return self._orbitals()[2]

    @property
    def energy_weighted_density(self) -> np.ndarray:
        """Canonical-orbital W = sum_s C_o eps_o C_o^T (aimd/qc/scf.py:935-941); the forces use
        sum_s D_s F_s D_s instead (gradients.energy_weighted_density)."""
        return self._orbitals()[3]

    @property
    def unrestricted(self) -> bool:
        return self.reference == "uhf"

    @property
    def density(self) -> np.ndarray:
        return self.spin_densities[0] + self.spin_densities[1]

    @property
    def density_alpha(self) -> np.ndarray:
        return self.spin_densities[0]

    @property
    def density_beta(self) -> np.ndarray:
        return self.spin_densities[1]

    @property
    def spin_density(self) -> np.ndarray:
        return self.spin_densities[0] - self.spin_densities[1]

    @property
    def guess_density(self) -> np.ndarray:
        return self.spin_densities.copy() if self.unrestricted else self.density

    @property
    def nmo(self) -> int:
        return int(self.mo_coeff.shape[-1])

    @property
    def s2_exact(self) -> float:
        sz = 0.5 * (self.n_alpha - self.n_beta)
        return sz * (sz + 1.0)

    def coulomb_exchange_densities(self) -> tuple[np.ndarray, list[np.ndarray], float]:
        if self.unrestricted:
            return self.density, [self.density_alpha, self.density_beta], 1.0
        P = self.density
        return P, [P], 0.5

    def __repr__(self) -> str:
        flag = "converged" if self.converged else "NOT CONVERGED"
        return (f"SCFResult({self.reference.upper()}/{self.basis.name}, E = {self.energy:.10f} Eh, "
                f"{self.iterations} iterations, {flag})")


# ================================================================= solver
class SCFSolver:
    """
    Reusable RHF/UHF solver for one molecule (aimd/qc/scf.py:687-965): the
    basis is parsed and staged on the device once; every run moves it to new
    positions and iterates from the requested guess.
    """

    def __init__(self, symbols: Sequence[str], basis: str | BasisSet = "sto-3g", charge: int = 0,
                 multiplicity: int = 1, reference: str | None = None, options: SCFOptions | None = None,
                 **overrides: Any) -> None:
        opt = options or SCFOptions()
        if overrides:
            known = {f.name for f in fields(SCFOptions)}
            bad = sorted(set(overrides) - known)
            if bad:
                raise TypeError(f"unknown SCF option(s): {', '.join(bad)}")
            opt = replace(opt, **overrides)
        self.options = opt
        n = len(symbols)
        if isinstance(basis, BasisSet):
            if [normalize_symbol(s) for s in symbols] != basis.symbols:
                raise ValueError("symbols do not match the atoms of the given BasisSet")
            self.basis = basis
        else:
            self.basis = build_basis(symbols, np.zeros((n, 3)), basis)
        self.symbols = list(self.basis.symbols)
        self.charge = int(charge)
        self.multiplicity = int(multiplicity)
        if reference is None:
            ref = "rhf" if self.multiplicity == 1 else "uhf"
        else:
            key = str(reference).strip().lower()
            if key not in _REFERENCES:
                raise ValueError(f"unknown reference {reference!r}; use 'rhf' or 'uhf'")
            ref = _REFERENCES[key]
        self.reference = ref
        self.n_alpha, self.n_beta = electron_counts(
            int(self.basis.atomic_numbers.sum()), self.charge, self.multiplicity, ref)
        if self.n_alpha > self.basis.nao:
            raise ValueError(f"{self.n_alpha} alpha electrons do not fit in {self.basis.nao} AOs")
        self.engine = qengine.Engine(self.basis)
        self.engine.set_jk_mode(self.options.jk_mode, self.options.max_store_bytes)
        self.engine.set_precision(self.options.precision, self.options.jk_fp32_switch,
                                  self.options.grad_fp32_switch, self.options.delta_fp32_switch)
        self._ints: SCFIntegrals | None = None
        self._ints_positions: np.ndarray | None = None
        self._ints_options: SCFOptions | None = None
        self._sad: np.ndarray | None = None

    @property
    def unrestricted(self) -> bool:
        return self.reference == "uhf"

    @property
    def nao(self) -> int:
        return self.basis.nao

    @property
    def nelectron(self) -> int:
        return self.n_alpha + self.n_beta

    @property
    def density_shape(self) -> tuple[int, ...]:
        n = self.nao
        return (2, n, n) if self.unrestricted else (n, n)

    @property
    def device(self) -> torch.device:
        return self.engine.device

    def integrals(self, positions: np.ndarray) -> SCFIntegrals:
        """Integrals at ``positions`` (bohr); cached for the last geometry (aimd/qc/scf.py:770-782)."""
        x = np.array(positions, dtype=float).reshape(-1, 3)
        if x.shape[0] != self.basis.natm:
            raise ValueError(f"expected {self.basis.natm} positions, got {x.shape[0]}")
        check_nuclear_separation(x, self.basis.atomic_numbers)
        if (self._ints is None or self._ints_options is not self.options
                or not np.array_equal(x, self._ints_positions)):
            self._ints = None
            self._ints = build_integrals(self.basis.with_positions(x), self.engine, self.options)
            self._ints_positions = x
            self._ints_options = self.options
        return self._ints

    def clear_cache(self) -> None:
        self._ints = None
        self._ints_positions = None

    def close(self) -> None:
        self.clear_cache()
        self.engine.close()

    # ---------------------------------------------------------------- guesses
    def _nocc(self) -> tuple[int, ...]:
        return (self.n_alpha,) if not self.unrestricted else (self.n_alpha, self.n_beta)

    def _occupied_density(self, F: torch.Tensor, X: torch.Tensor, nocc: int) -> torch.Tensor:
        _, Cp = torch.linalg.eigh(X.T @ F @ X)
        C = X @ Cp[:, :nocc]
        return C @ C.T

    def initial_spin_densities(self, I: SCFIntegrals, guess: Any) -> tuple[torch.Tensor, str]:
        """(nspin, nao, nao) starting spin densities (aimd/qc/scf.py:793-827)."""
        nspin = 2 if self.unrestricted else 1
        n = self.nao
        dev = self.device
        if isinstance(guess, SCFResult):
            if guess.basis.nao != n:
                raise ValueError("guess SCFResult has a different basis")
            D = guess.spin_densities
            Ds = D[:nspin].copy() if self.unrestricted else 0.5 * (D[0] + D[1])[None]
            return torch.as_tensor(Ds, dtype=torch.float64, device=dev), "result"
        if guess is None:
            guess = self.options.guess
        if isinstance(guess, str):
            kind = guess.strip().lower()
            if kind == "sad":
                if self._sad is None:
                    self._sad = sad_density(self.basis)
                Ds = np.repeat(0.5 * self._sad[None], nspin, axis=0)
                return torch.as_tensor(Ds, dtype=torch.float64, device=dev), kind
            if kind in ("core", "gwh"):
                if kind == "core":
                    F = I.H
                else:
                    h = torch.diagonal(I.H)
                    F = 0.5 * GWH_K * I.S * (h[:, None] + h[None, :])
                    F = F - torch.diag(torch.diagonal(F)) + torch.diag(h)
                return torch.stack([self._occupied_density(F, I.X, k) for k in self._nocc()]), kind
            raise ValueError(f"unknown guess {guess!r}; use one of {_GUESSES} or a density")
        if isinstance(guess, torch.Tensor):
            # a device density (e.g. an XL-BOMD auxiliary density): no host round trip
            Dt = guess.to(device=dev, dtype=torch.float64)
            if tuple(Dt.shape) == (n, n):
                Dt = (0.5 * Dt)[None].expand(nspin, n, n)
            elif tuple(Dt.shape) == (2, n, n):
                Dt = Dt if self.unrestricted else (0.5 * (Dt[0] + Dt[1]))[None]
            else:
                raise ValueError(
                    f"density guess has shape {tuple(Dt.shape)}; expected {self.density_shape} "
                    f"({self.reference.upper()}) or {(n, n) if self.unrestricted else (2, n, n)}")
            if not bool(torch.isfinite(Dt).all()):
                raise ValueError("density guess contains non-finite values")
            return (0.5 * (Dt + Dt.transpose(1, 2))).contiguous(), "density"
        D = np.array(guess, dtype=float)
        if D.shape == (n, n):
            Ds = np.repeat(0.5 * D[None], nspin, axis=0)
        elif D.shape == (2, n, n):
            Ds = D.copy() if self.unrestricted else 0.5 * (D[0] + D[1])[None]
        else:
            raise ValueError(
                f"density guess has shape {D.shape}; expected {self.density_shape} "
                f"({self.reference.upper()}) or {(n, n) if self.unrestricted else (2, n, n)}")
        if not np.all(np.isfinite(Ds)):
            raise ValueError("density guess contains non-finite values")
        Ds = 0.5 * (Ds + Ds.transpose(0, 2, 1))
        return torch.as_tensor(Ds, dtype=torch.float64, device=dev), "density"

    # ------------------------------------------------------------------- SCF
    def run(self, positions: np.ndarray, guess: Any = None) -> SCFResult:
        return self._run(positions, guess, stacklevel=3)

    def _run(self, positions: np.ndarray, guess: Any, stacklevel: int) -> SCFResult:
        t0 = time.perf_counter()
        I = self.integrals(positions)
        torch.cuda.synchronize()
        t1 = time.perf_counter()
        Ds, label = self.initial_spin_densities(I, guess)
        t2 = time.perf_counter()
        if I.nmo < self.n_alpha:
            raise ValueError(f"{self.n_alpha} alpha electrons do not fit in {I.nmo} "
                             "linearly independent orbitals")
        state = self._iterate(I, Ds)
        torch.cuda.synchronize()
        t3 = time.perf_counter()
        res = self._result(I, state, label)
        res.timings = {"integrals": t1 - t0, "guess": t2 - t1, "iterations": t3 - t2,
                       "total": time.perf_counter() - t0}
        if not res.converged and self.options.warn_unconverged:
            warnings.warn(
                f"SCF not converged after {res.iterations} iterations "
                f"(dE = {res.energy_change:.2e} Eh, commutator {res.commutator_norm:.2e}); "
                "returning the last iterate flagged converged=False",
                SCFConvergenceWarning, stacklevel=stacklevel)
        return res

    def _fock(self, I: SCFIntegrals, Ds: torch.Tensor, jk: tuple | None = None) -> torch.Tensor:
        g = 2.0 / Ds.shape[0]
        J, K = I.engine.jk(Ds) if jk is None else jk
        F = torch.add(I.H, J[0] if J.shape[0] == 1 else J.sum(dim=0), alpha=g) - K
        return 0.5 * (F + F.transpose(1, 2))

    @property
    def incremental_threshold(self) -> float:
        """Screening threshold of the density-difference builds (SCFOptions docstring)."""
        opt = self.options
        if opt.incremental_threshold is not None:
            return opt.incremental_threshold
        if opt.precision == "mixed":
            return max(opt.schwarz_threshold, 0.1 * opt.conv_energy)
        return opt.schwarz_threshold

    @property
    def incremental(self) -> bool:
        """Whether SCF iterations after the first build J/K from density differences."""
        inc = self.options.incremental_fock
        return self.options.precision == "mixed" if inc is None else bool(inc)

    def _next_density(self, Fo: torch.Tensor, X: torch.Tensor, nocc: tuple[int, ...],
                      Ds: torch.Tensor) -> torch.Tensor:
        """Aufbau densities of the orthogonal-basis Fock matrices Fo (nspin, nmo, nmo)."""
        Dn = torch.empty_like(Ds)
        if self.options.density_solver == "purify":
            purify = purified_projector_device if Fo.is_cuda else purified_projector
            for s in range(Fo.shape[0]):
                P = purify(Fo[s], nocc[s])
                if P is None:
                    break
                torch.matmul(X @ P, X.T, out=Dn[s])
            else:
                return Dn
        _, Cp = torch.linalg.eigh(Fo)
        for s in range(Fo.shape[0]):
            C = X @ Cp[s][:, :nocc[s]]
            Dn[s] = C @ C.T
        return Dn

    def _iterate(self, I: SCFIntegrals, Ds: torch.Tensor) -> dict[str, Any]:
        opt = self.options
        # every SCF starts the purification replay counts afresh, so a result depends
        # on its own inputs only (bit-for-bit restarts), not on earlier SCFs
        for g in _TC2_GRAPHS.values():
            g.needed = 1
        nspin = Ds.shape[0]
        g = 2.0 / nspin
        nocc = self._nocc()
        S, H, X = I.S, I.H, I.X
        XtS = X.T @ S
        diis = DIIS(opt.diis_space) if opt.diis else None
        e_prev = np.nan
        best = np.inf
        since_best = 0
        restarts = 0
        history: list[dict[str, float]] = []
        converged = False
        eye = torch.eye(I.nmo, dtype=torch.float64, device=S.device)
        incremental = self.incremental
        dthr = self.incremental_threshold
        jk = None
        D_built = None
        for it in range(1, opt.max_iter + 1):
            if incremental:
                # J(D) = J(D_built) + J(D - D_built): the first build is complete
                if jk is None:
                    jk = I.engine.jk(Ds)
                else:
                    dJ, dK = I.engine.jk_delta(Ds - D_built, dthr)
                    jk = (jk[0] + dJ, jk[1] + dK)
                D_built = Ds
            F = self._fock(I, Ds, jk)
            e_dev = 0.5 * g * torch.vdot(Ds.reshape(-1), (H + F).reshape(-1))
            FDS = F @ Ds @ S
            err = X.T @ (FDS - FDS.transpose(1, 2)) @ X
            Fo = X.T @ F @ X
            # the DIIS vector is pushed before the convergence test, so its Gram
            # matrix comes back with the energy and the commutator in one read
            # max|err| from its extremes (one reduction kernel; torch's inf-norm takes
            # several times longer on this size)
            mn, mx = torch.aminmax(err)
            parts = [e_dev.reshape(1), mx.reshape(1), mn.reshape(1)]
            n_diis = len(diis) if diis is not None else 0
            if diis is not None:
                diis.push(Fo, err)
                parts.append(diis.raw_gram.reshape(-1))
            scal = torch.cat(parts).cpu().numpy()
            e_el, err_norm = float(scal[0]), max(float(scal[1]), -float(scal[2]))
            de = e_el - e_prev
            history.append({"energy": e_el + I.nuclear_repulsion, "energy_change": de,
                            "commutator": err_norm, "diis": n_diis})
            if abs(de) < opt.conv_energy and err_norm < opt.conv_commutator:
                converged = True
                break
            if it == opt.max_iter:
                break
            if diis is not None:
                Gr = diis.gram_from(scal[3:].reshape(diis.space, diis.space))
                if err_norm < best:
                    best, since_best = err_norm, 0
                else:
                    since_best += 1
                if opt.diis_restart and since_best >= opt.diis_restart:
                    diis.keep_last()
                    Gr = None
                    restarts += 1
                    best, since_best = err_norm, 0
                Fo = diis.extrapolate(Gr)
            stabilize = err_norm > opt.stabilize_until
            if stabilize and opt.level_shift > 0.0:
                Do = XtS @ Ds @ XtS.T
                Fo = Fo + opt.level_shift * (eye - Do)
            Dn = self._next_density(Fo, X, nocc, Ds)
            if stabilize and opt.damping > 0.0:
                Dn = (1.0 - opt.damping) * Dn + opt.damping * Ds
            Ds = Dn
            e_prev = e_el
        return {"Ds": Ds, "F": F, "e_el": e_el, "de": de, "err": err_norm, "it": it,
                "converged": converged, "history": history, "restarts": restarts}

    def _result(self, I: SCFIntegrals, st: dict[str, Any], label: str) -> SCFResult:
        basis = I.basis
        Ds, F = st["Ds"], st["F"]
        nspin = Ds.shape[0]
        g = 2.0 / nspin
        nocc = self._nocc()
        X, S = I.X, I.S
        nmo = I.nmo

        def orbitals():
            eps_t, Cp = torch.linalg.eigh(X.T @ F @ X)
            C_t = X @ Cp
            W = torch.zeros((I.nao, I.nao), dtype=torch.float64, device=S.device)
            for s in range(nspin):
                Co = C_t[s][:, :nocc[s]]
                W = W + g * (Co * eps_t[s, :nocc[s]]) @ Co.T
            eps, C, Wh = eps_t.cpu().numpy(), C_t.cpu().numpy(), (0.5 * (W + W.T)).cpu().numpy()
            occ = np.zeros((nspin, nmo))
            for s in range(nspin):
                occ[s, :nocc[s]] = g
            if nspin == 1:
                eps, C, occ = eps[0], C[0], occ[0]
            return eps, C, occ, Wh

        spin_t = torch.stack([Ds[0], Ds[-1]])
        P = spin_t[0] + spin_t[1]
        PS_diag = torch.diagonal(P @ S)
        dip_el = -(I.M * P).sum(dim=(1, 2))
        s2_t = torch.sum((spin_t[0] @ S) * (spin_t[1] @ S).T) if nspin == 2 else None
        host = [spin_t, F, PS_diag, dip_el]
        if s2_t is not None:
            host.append(s2_t)
        host = [t.cpu().numpy() for t in host]
        spin, Fh, psd, dipel = host[:4]
        q = basis.nuclear_charges.copy()
        np.subtract.at(q, basis.ao_atom, psd)
        dip = dipel + nuclear_dipole(basis.nuclear_charges, basis.positions)
        sz = 0.5 * (self.n_alpha - self.n_beta)
        s2 = sz * (sz + 1.0) + self.n_beta - float(host[4]) if nspin == 2 else 0.0
        if nspin == 1:
            Fh = Fh[0]
        e_el = st["e_el"]
        return SCFResult(
            reference=self.reference, converged=st["converged"], iterations=st["it"],
            energy=e_el + I.nuclear_repulsion, electronic_energy=e_el,
            nuclear_repulsion=I.nuclear_repulsion,
            spin_densities=spin, fock=Fh,
            mulliken_charges=q, dipole=dip, s2=float(s2), n_alpha=self.n_alpha,
            n_beta=self.n_beta, energy_change=float(st["de"]), commutator_norm=st["err"],
            basis=basis, n_removed=I.n_removed, guess=label, diis_restarts=st["restarts"],
            history=st["history"], integrals=I, device_spin_densities=Ds, device_fock=F, orbitals=orbitals)


def run_scf(symbols: Sequence[str], positions: np.ndarray, basis: str | BasisSet = "sto-3g", charge: int = 0,
            multiplicity: int = 1, reference: str | None = None, guess: Any = None,
            options: SCFOptions | None = None, **overrides: Any) -> SCFResult:
    """One-shot SCF: ``SCFSolver(...).run(positions, guess)`` (aimd/qc/scf.py:968-982)."""
    solver = SCFSolver(symbols, basis, charge=charge, multiplicity=multiplicity,
                       reference=reference, options=options, **overrides)
    return solver._run(positions, guess, stacklevel=3)
