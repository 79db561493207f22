"""
``hf`` force backend on MI355X: RHF/UHF energies and analytic gradients from
the gfx950 engine, a drop-in for the reference's ``hf`` backend
(aimd/backends/hf.py:150-369): same constructor arguments and validation,
same density-guess protocol (external > previous > init), same ``info`` keys
and error behaviour (ValueError for bad input or coincident nuclei;
non-convergence flagged, never raised).

Per ``compute``: the positions go to the device plan (pair data, Schwarz
bounds, screened quartet list), the SCF iterates with integral-direct J/K on
the GPU, and the gradient is the backward() of the energy's
torch.autograd.Function (derivative kernels contracted with Gamma and W).

Extra keyword arguments beyond the reference's: ``device`` (torch device
index or string; default the current HIP device).
"""

from __future__ import annotations

import os
import time
from typing import Any, Sequence

import numpy as np
import torch

from aimd_amd.backends.base import ForceBackend, GradientResult
from aimd_amd.backends.registry import register_backend
from aimd_amd.qc import engine as qengine
from aimd_amd.qc.autograd import energy_and_forces
from aimd_amd.qc.basis import BasisSet
from aimd_amd.qc.scf import SCFOptions, SCFResult, SCFSolver

_REFERENCES = {"auto": None, "rhf": "rhf", "restricted": "rhf", "uhf": "uhf", "unrestricted": "uhf"}
_METHODS = {"hf": None, "scf": None, "rhf": "rhf", "uhf": "uhf"}
_GUESSES = {"auto": "sad", "sad": "sad", "core": "core", "gwh": "gwh"}
_OWN_OPTIONS = {"conv_energy": "conv_tol", "conv_commutator": "conv_tol_grad",
                "max_iter": "max_cycles", "guess": "guess", "warn_unconverged": None}


@register_backend
class HFBackend(ForceBackend):
    """Native RHF / UHF energies and analytic gradients on the GPU (module docstring)."""

    name = "hf"
    description = "native RHF / UHF, analytic gradients (McMurchie-Davidson on MI355X)"
    supports_density_guess = True

    def __init__(
        self,
        symbols: Sequence[str],
        charge: int = 0,
        multiplicity: int = 1,
        basis: str | BasisSet = "sto-3g",
        reference: str | None = "auto",
        method: str = "hf",
        conv_tol: float = 1e-10,
        conv_tol_grad: float = 1e-7,
        max_cycles: int = 100,
        guess: str = "auto",
        reuse_density: bool = True,
        gradient_screening: float = qengine.DEFAULT_SCHWARZ_GRAD,
        threads: int | None = None,
        scf_options: dict[str, Any] | None = None,
        device: int | str | None = None,
    ) -> None:
        super().__init__(symbols, charge, multiplicity)
        ref = self._parse_reference(reference, method)
        key = str(guess).strip().lower()
        if key not in _GUESSES:
            raise ValueError(f"unknown guess {guess!r}; use one of {sorted(_GUESSES)}")
        self.init_guess = _GUESSES[key]
        if not gradient_screening >= 0.0:
            raise ValueError("gradient_screening must be >= 0")
        self.gradient_screening = float(gradient_screening)
        self.reuse_density = bool(reuse_density)
        extra = dict(scf_options or {})
        for k in extra:
            if k in _OWN_OPTIONS:
                arg = _OWN_OPTIONS[k]
                raise ValueError(f"set SCF option {k!r} through the backend argument {arg!r}"
                                 if arg else f"SCF option {k!r} is managed by the backend")
        try:
            options = SCFOptions(conv_energy=float(conv_tol), conv_commutator=float(conv_tol_grad),
                                 max_iter=int(max_cycles), guess=self.init_guess,
                                 warn_unconverged=False, **extra)
        except TypeError as e:
            raise ValueError(f"invalid scf_options: {e}") from e
        ncpu = os.cpu_count() or 1
        if threads is not None and not 1 <= int(threads) <= ncpu:
            raise ValueError(f"threads must be in 1..{ncpu}")
        self.threads = None if threads is None else int(threads)
        if device is not None:
            torch.cuda.set_device(torch.device(device) if isinstance(device, str) else int(device))
        self.solver = SCFSolver(self.symbols, basis, charge=self.charge, multiplicity=self.multiplicity,
                                reference=ref, options=options)
        self._guess: np.ndarray | None = None
        self._last_density: np.ndarray | None = None
        self.last_scf: SCFResult | None = None

    def _parse_reference(self, reference: str | None, method: str) -> str | None:
        m = str(method).strip().lower()
        if m not in _METHODS:
            raise ValueError(f"the hf backend does Hartree-Fock only; unknown method {method!r} "
                             "(use 'hf', 'rhf' or 'uhf', or the pyscf backend for DFT / MP2)")
        r = "auto" if reference is None else str(reference).strip().lower()
        if r not in _REFERENCES:
            raise ValueError(f"unknown reference {reference!r}; use 'auto', 'rhf' or 'uhf'")
        ref, forced = _REFERENCES[r], _METHODS[m]
        if ref is not None and forced is not None and ref != forced:
            raise ValueError(f"method {method!r} conflicts with reference {reference!r}")
        return ref or forced

    @property
    def options(self) -> SCFOptions:
        return self.solver.options

    @property
    def unrestricted(self) -> bool:
        return self.solver.unrestricted

    @property
    def nao(self) -> int:
        return self.solver.nao

    @property
    def basis(self) -> BasisSet:
        return self.solver.basis

    @property
    def density_shape(self) -> tuple[int, ...]:
        return self.solver.density_shape

    @property
    def label(self) -> str:
        return "UHF" if self.unrestricted else "RHF"

    def set_density_guess(self, density) -> None:
        if isinstance(density, torch.Tensor):
            if tuple(density.shape) != self.density_shape:
                raise ValueError(
                    f"density guess has shape {tuple(density.shape)}, expected {self.density_shape} "
                    f"({'unrestricted' if self.unrestricted else 'restricted'} reference)")
            if not bool(torch.isfinite(density).all()):
                raise ValueError("density guess contains non-finite values")
            self._guess = density.detach().to(self.solver.device, torch.float64).clone()
            return
        d = np.array(density, dtype=float)
        if d.shape != self.density_shape:
            raise ValueError(
                f"density guess has shape {d.shape}, expected {self.density_shape} "
                f"({'unrestricted' if self.unrestricted else 'restricted'} reference)")
        if not np.all(np.isfinite(d)):
            raise ValueError("density guess contains non-finite values")
        self._guess = d

    def device_density(self) -> torch.Tensor:
        """The last SCF density on the device, in the density-guess layout (RHF: P (nao,
        nao); UHF: (2, nao, nao)); what XL-BOMD propagates without a host round trip."""
        res = self.last_scf
        if res is None or res.device_spin_densities is None:
            raise RuntimeError("no SCF has run yet")
        Ds = res.device_spin_densities
        if self.unrestricted:
            return Ds.clone()
        return Ds[0] + Ds[-1] if Ds.shape[0] == 2 else 2.0 * Ds[0]

    def reset_guess(self) -> None:
        self._guess = None
        self._last_density = None

    def compute(self, positions: np.ndarray) -> GradientResult:
        t0 = time.perf_counter()
        x = np.array(positions, dtype=float)
        if x.ndim != 2 or x.shape != (len(self.symbols), 3):
            raise ValueError(f"expected positions of shape ({len(self.symbols)}, 3), got {x.shape}")
        if self._guess is not None:
            guess, label = self._guess, "external"
        elif self.reuse_density and self._last_density is not None:
            guess, label = self._last_density, "previous"
        else:
            guess, label = self.init_guess, "init"
        self._guess = None
        energy, gradient, res = energy_and_forces(self.solver, x, guess, self.gradient_screening)
        t2 = time.perf_counter()
        t1 = t0 + res.timings.get("total", 0.0)
        density = res.guess_density
        if self.reuse_density:
            self._last_density = density
        self.last_scf = res
        warnings_: list[str] = []
        if not res.converged:
            warnings_.append(
                f"SCF not converged after {res.iterations} iterations (energy change "
                f"{res.energy_change:.2e} Eh, commutator {res.commutator_norm:.2e}); "
                "energy and gradient are those of the last iterate")
        if res.n_removed:
            warnings_.append(
                f"{res.n_removed} near-linearly-dependent basis combination(s) removed; "
                "the analytic gradient is then not the exact derivative of the energy")
        info: dict[str, Any] = {
            "method": self.label,
            "basis": res.basis.name,
            "reference": res.reference,
            "scf_converged": bool(res.converged),
            "scf_iterations": int(res.iterations),
            "scf_energy": float(res.energy),
            "electronic_energy": float(res.electronic_energy),
            "nuclear_repulsion": float(res.nuclear_repulsion),
            "energy_change": float(res.energy_change),
            "commutator_norm": float(res.commutator_norm),
            "s2": float(res.s2),
            "s2_exact": float(res.s2_exact),
            "mulliken_charges": res.mulliken_charges.copy(),
            "guess": label,
            "init_guess": self.init_guess,
            "n_removed": int(res.n_removed),
            "warnings": warnings_,
            "timings": {
                "integrals": float(res.timings.get("integrals", 0.0)),
                "guess": float(res.timings.get("guess", 0.0)),
                "scf_iterations": float(res.timings.get("iterations", 0.0)),
                "scf": t1 - t0,
                "gradient": t2 - t1,
                "total": time.perf_counter() - t0,
            },
        }
        return GradientResult(energy=float(energy), gradient=gradient, converged=bool(res.converged),
                              density=density.copy(), dipole=np.array(res.dipole, dtype=float), info=info)

    def close(self) -> None:
        """Release the device state of the last geometry and forget stored densities."""
        self.solver.clear_cache()
        self.reset_guess()
        self.last_scf = None
