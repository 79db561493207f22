"""
Analytic nuclear gradient of an RHF/UHF energy (aimd/qc/gradients.py).

    dE/dR = dE_nuc/dR + sum P d(T+V)/dR - sum W dS/dR + 1/2 sum Gamma d(mn|ls)/dR,
    W = sum_s D_s F_s D_s

formed from the density and Fock matrix of the SCF's last iteration (the
convention of aimd/qc/gradients.py:106-125, 168-193). The one-electron and
derivative-ERI contractions run on the gfx950 engine; the derivative
integrals are recomputed on the fly and never stored.
"""

from __future__ import annotations

import time
from dataclasses import dataclass, field

import numpy as np
import torch

from aimd_amd.qc import engine as qengine
from aimd_amd.qc.scf import SCFResult, nuclear_repulsion_gradient


def energy_weighted_density(spin_densities, fock):
    """W = sum_s D_s F_s D_s, symmetrised (numpy or torch; aimd/qc/gradients.py:106-125)."""
    if isinstance(spin_densities, torch.Tensor):
        D, F = spin_densities, fock
        if F.dim() == 2:
            F = F.expand(2, *F.shape)
        W = D[0] @ F[0] @ D[0] + D[1] @ F[1] @ D[1]
        return 0.5 * (W + W.T)
    D = np.asarray(spin_densities, dtype=float)
    F = np.asarray(fock, dtype=float)
    if D.ndim != 3 or D.shape[0] != 2 or D.shape[1] != D.shape[2]:
        raise ValueError(f"spin_densities must have shape (2, nao, nao), got {D.shape}")
    n = D.shape[1]
    if F.shape == (n, n):
        W = D[0] @ F @ D[0] + D[1] @ F @ D[1]
    elif F.shape == (2, n, n):
        W = D[0] @ F[0] @ D[0] + D[1] @ F[1] @ D[1]
    else:
        raise ValueError(f"fock must have shape ({n}, {n}) or (2, {n}, {n}), got {F.shape}")
    return 0.5 * (W + W.T)


@dataclass(frozen=True)
class GradientTerms:
    """
    Parts of an SCF gradient, (natm, 3) each, Eh/bohr (aimd/qc/gradients.py:128-156):
    nuclear_repulsion, kinetic (sum P dT/dR), nuclear_attraction (sum P dV/dR,
    basis-function and operator centres), overlap (-sum W dS/dR), two_electron
    (1/2 sum Gamma d(mn|ls)/dR).
    """
    nuclear_repulsion: np.ndarray
    kinetic: np.ndarray
    nuclear_attraction: np.ndarray
    overlap: np.ndarray
    two_electron: np.ndarray
    timings: dict = field(default_factory=dict)

    @property
    def one_electron(self) -> np.ndarray:
        return self.kinetic + self.nuclear_attraction + self.overlap

    @property
    def total(self) -> np.ndarray:
        return self.nuclear_repulsion + self.one_electron + self.two_electron


def _device_state(result: SCFResult):
    if result.integrals is None or result.device_spin_densities is None:
        raise ValueError("the SCFResult carries no device state (it was not produced by this engine)")
    Ds = result.device_spin_densities
    F = result.device_fock
    spin = torch.stack([Ds[0], Ds[-1]])
    Fs = torch.stack([F[0], F[-1]])
    return result.integrals, spin, Fs


def _check(result: SCFResult, require_converged: bool) -> None:
    if not isinstance(result, SCFResult):
        raise TypeError(f"expected an SCFResult, got {type(result).__name__}")
    if require_converged and not result.converged:
        raise ValueError(
            f"SCF not converged ({result.iterations} iterations, commutator "
            f"{result.commutator_norm:.2e}); its analytic gradient is not dE/dR")


def _two_electron(I, result: SCFResult, spin: torch.Tensor, P: torch.Tensor, threshold: float) -> torch.Tensor:
    if result.unrestricted:
        return I.engine.grad_two(P, spin, 1.0, threshold)
    return I.engine.grad_two(P, P[None], 0.5, threshold)


def gradient_terms(result: SCFResult, schwarz_threshold: float = qengine.DEFAULT_SCHWARZ_GRAD,
                   require_converged: bool = False, blas_threads=None) -> GradientTerms:
    """The gradient of ``result`` split into its physical parts (aimd/qc/gradients.py:196-223);
    the one-electron terms come from one engine call with three kernel passes."""
    _check(result, require_converged)
    t0 = time.perf_counter()
    I, spin, Fs = _device_state(result)
    b = I.basis
    W = energy_weighted_density(spin, Fs)
    P = spin[0] + spin[1]
    g_nuc = nuclear_repulsion_gradient(b.nuclear_charges, b.positions)
    t1 = time.perf_counter()
    gt, gv, gs = I.engine.grad_one_terms(P, W)
    torch.cuda.synchronize()
    t2 = time.perf_counter()
    g2 = _two_electron(I, result, spin, P, schwarz_threshold)
    parts = [t.cpu().numpy() for t in (gt, gv, gs, g2)]
    t3 = time.perf_counter()
    timings = {"nuclear_and_w": t1 - t0, "one_electron": t2 - t1, "two_electron": t3 - t2, "total": t3 - t0}
    return GradientTerms(g_nuc, *parts, timings=timings)


def scf_gradient(result: SCFResult, schwarz_threshold: float = qengine.DEFAULT_SCHWARZ_GRAD,
                 require_converged: bool = False, blas_threads=None) -> np.ndarray:
    """Analytic dE/dR (hartree/bohr, (natm, 3)) of an SCF result (aimd/qc/gradients.py:168-193):
    one fused one-electron pass and one derivative-ERI pass on the device."""
    _check(result, require_converged)
    I, spin, Fs = _device_state(result)
    b = I.basis
    W = energy_weighted_density(spin, Fs)
    P = spin[0] + spin[1]
    g1 = I.engine.grad_one(P, W)
    g2 = _two_electron(I, result, spin, P, schwarz_threshold)
    g = (g1 + g2).cpu().numpy()
    return nuclear_repulsion_gradient(b.nuclear_charges, b.positions) + g


def invariance_residuals(positions: np.ndarray, gradient: np.ndarray) -> tuple[np.ndarray, np.ndarray]:
    """(net force, net torque about the centroid) of a gradient (aimd/qc/gradients.py:226-238)."""
    x = np.asarray(positions, dtype=float).reshape(-1, 3)
    g = np.asarray(gradient, dtype=float).reshape(-1, 3)
    if x.shape != g.shape:
        raise ValueError(f"positions {x.shape} and gradient {g.shape} differ in shape")
    return g.sum(axis=0), np.cross(x - x.mean(axis=0), g).sum(axis=0)
