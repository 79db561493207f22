"""
Force-provider interface between an MD loop and an electronic-structure code
(the reference's plugin API, aimd/backends/base.py:42-84, kept identical so
the engine is a drop-in behind it).

Unit contract: ``compute`` receives positions in bohr, shape (N, 3), and
returns the energy in hartree and the gradient dE/dR in hartree/bohr, shape
(N, 3). Forces are ``-gradient``.

Density-guess protocol: a backend with ``supports_density_guess = True``
returns its final AO density in ``GradientResult.density`` -- (nao, nao) total
density for restricted references, (2, nao, nao) [alpha, beta] for
unrestricted ones -- and accepts an array of that shape in
``set_density_guess``, used as the start of the next ``compute`` only.
"""

from __future__ import annotations

from abc import ABC, abstractmethod
from dataclasses import dataclass, field
from typing import Any, ClassVar, Sequence

import numpy as np


@dataclass
class GradientResult:
    """Energy and nuclear gradient for one geometry."""
    energy: float
    gradient: np.ndarray
    converged: bool = True
    density: np.ndarray | None = None
    dipole: np.ndarray | None = None
    info: dict[str, Any] = field(default_factory=dict)

    @property
    def forces(self) -> np.ndarray:
        return -self.gradient


class ForceBackend(ABC):
    """Abstract energy/gradient provider."""

    name: ClassVar[str] = ""
    supports_density_guess: ClassVar[bool] = False

    def __init__(self, symbols: Sequence[str], charge: int = 0, multiplicity: int = 1) -> None:
        self.symbols = list(symbols)
        self.charge = int(charge)
        self.multiplicity = int(multiplicity)

    @abstractmethod
    def compute(self, positions: np.ndarray) -> GradientResult:
        """Energy and gradient at ``positions`` (bohr, shape (N, 3))."""

    def set_density_guess(self, density: np.ndarray) -> None:
        raise NotImplementedError(f"backend '{self.name}' does not support density guesses")

    def close(self) -> None:
        """Release resources (device memory). Optional."""
