"""
Minimal molecular-dynamics driver around a force backend: Maxwell-Boltzmann
velocities and velocity-Verlet NVE steps, with the semantics of the
reference's caller (aimd/system.py:234-268, aimd/integrators.py:215-232), so
that benchmark and drift measurements step the engine exactly as the
reference's MD loop would.
"""

from __future__ import annotations

import numpy as np

from aimd_amd.elements import ATOMIC_MASSES_AMU, normalize_symbol
from aimd_amd.units import AMU_TO_AU, FS_TO_AU_TIME, KB_AU


def masses_au(symbols) -> np.ndarray:
    return np.array([ATOMIC_MASSES_AMU[normalize_symbol(s)] for s in symbols]) * AMU_TO_AU


def kinetic_energy(masses: np.ndarray, v: np.ndarray) -> float:
    return 0.5 * float(np.sum(masses[:, None] * v * v))


def temperature(masses: np.ndarray, v: np.ndarray, com_removed: bool = True) -> float:
    ndof = 3 * len(masses) - (3 if com_removed else 0)
    return 2.0 * kinetic_energy(masses, v) / (ndof * KB_AU) if ndof > 0 else 0.0


def maxwell_boltzmann(masses: np.ndarray, temperature_k: float, seed=None) -> np.ndarray:
    """Velocities at ``temperature_k``: normal draws, COM motion removed, rescaled exactly
    (MolecularSystem.initialize_velocities with exact=True, remove_rotation=False)."""
    rng = np.random.default_rng(seed)
    sigma = np.sqrt(KB_AU * temperature_k / masses)
    v = rng.normal(size=(len(masses), 3)) * sigma[:, None]
    v -= (masses[:, None] * v).sum(axis=0) / masses.sum()
    t = temperature(masses, v)
    if t > 0.0:
        v *= np.sqrt(temperature_k / t)
    return v


class VelocityVerletNVE:
    """x(t+dt) = x + dt v + dt^2/2 a; v(t+dt) = v + dt/2 (a + a') with one compute per step."""

    def __init__(self, backend, masses: np.ndarray, positions: np.ndarray, velocities: np.ndarray,
                 timestep_fs: float) -> None:
        self.backend = backend
        self.m = np.asarray(masses, dtype=float)
        self.x = np.array(positions, dtype=float)
        self.v = np.array(velocities, dtype=float)
        self.dt = float(timestep_fs) * FS_TO_AU_TIME
        self.result = None

    def _accel(self) -> np.ndarray:
        return -self.result.gradient / self.m[:, None]

    def initialize(self):
        self.result = self.backend.compute(self.x)
        return self.result

    def step(self):
        if self.result is None:
            self.initialize()
        self.v += 0.5 * self.dt * self._accel()
        self.x += self.dt * self.v
        self.result = self.backend.compute(self.x)
        self.v += 0.5 * self.dt * self._accel()
        return self.result

    def total_energy(self) -> float:
        return float(self.result.energy) + kinetic_energy(self.m, self.v)


def read_xyz(path: str) -> tuple[list[str], np.ndarray]:
    """Symbols and positions (angstrom) of an XYZ file."""
    with open(path, "r", encoding="utf-8") as f:
        lines = f.read().splitlines()
    n = int(lines[0].split()[0])
    sym, pos = [], []
    for line in lines[2:2 + n]:
        p = line.split()
        sym.append(p[0])
        pos.append([float(v) for v in p[1:4]])
    return sym, np.array(pos)
