"""
Molecular-dynamics drivers around a force backend, with the semantics of the
reference's callers so that the engine is stepped exactly as the reference's
MD loop steps it:

* :class:`System` -- positions (bohr), velocities (bohr / au time), masses
  (m_e) and the COM flag that sets the degrees of freedom
  (aimd/system.py:234-268 for the Maxwell-Boltzmann draw);
* :class:`VelocityVerlet` (NVE), :class:`XLBOMD` (velocity Verlet whose SCF
  guesses come from the dissipative extended-Lagrangian auxiliary density,
  aimd/xlbomd.py:78-145, aimd/integrators.py:321-416) and
  :class:`LangevinBAOAB` (aimd/integrators.py:419-491): one force call per
  step at the new positions, ``state_dict`` / ``load_state_dict`` for exact
  continuation;
* :func:`save_checkpoint` / :func:`load_checkpoint` -- the reference's
  checkpoint file (aimd/checkpoint.py:123-149, 253-348): an ``.npz`` whose
  arrays are members and whose remaining state is JSON, with the backend
  fingerprint checked on restart.

With the MI355X backend the XL-BOMD history stays on the device: the SCF
density comes from ``backend.device_density()`` and the propagated guess goes
back through ``set_density_guess`` as a device tensor. Other backends get
numpy arrays, as the reference hands them.
"""

from __future__ import annotations

import json
import math
import os
import zipfile
from dataclasses import dataclass, field
from pathlib import Path
from typing import Any

import numpy as np

from aimd_amd.elements import ATOMIC_MASSES_AMU, normalize_symbol
from aimd_amd.units import AMU_TO_AU, FS_TO_AU_TIME, KB_AU

try:                                    # device-resident XL-BOMD history with the GPU backend
    import torch
except ImportError:                     # pragma: no cover
    torch = None


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


@dataclass
class System:
    """Nuclear state the integrators advance in place."""
    symbols: list
    positions: np.ndarray
    velocities: np.ndarray
    masses: np.ndarray
    com_removed: bool = True

    @classmethod
    def create(cls, symbols, positions, velocities=None, masses=None) -> "System":
        x = np.array(positions, dtype=float).reshape(-1, 3)
        v = np.zeros_like(x) if velocities is None else np.array(velocities, dtype=float).reshape(x.shape)
        m = masses_au(symbols) if masses is None else np.array(masses, dtype=float)
        return cls([normalize_symbol(s) for s in symbols], x, v, m)

    def kinetic_energy(self) -> float:
        return kinetic_energy(self.masses, self.velocities)

    def temperature(self) -> float:
        return temperature(self.masses, self.velocities, self.com_removed)

    def copy(self) -> "System":
        return System(list(self.symbols), self.positions.copy(), self.velocities.copy(), self.masses.copy(),
                      self.com_removed)


# ----------------------------------------------------------------- results
def _result_state(r) -> dict | None:
    if r is None:
        return None
    return {"energy": float(r.energy), "gradient": np.array(r.gradient, dtype=float), "converged": bool(r.converged),
            "density": None if r.density is None else np.array(r.density),
            "dipole": None if r.dipole is None else np.array(r.dipole),
            "info": {k: (v.copy() if isinstance(v, np.ndarray) else v) for k, v in r.info.items()}}


def _result_from_state(st):
    from aimd_amd.backends.base import GradientResult
    if st is None:
        return None
    return GradientResult(energy=float(st["energy"]), gradient=np.array(st["gradient"], dtype=float),
                          converged=bool(st["converged"]),
                          density=None if st["density"] is None else np.array(st["density"]),
                          dipole=None if st["dipole"] is None else np.array(st["dipole"]),
                          info=dict(st.get("info") or {}))


# -------------------------------------------------------------- integrators
class Integrator:
    """One backend call per step, at the new positions; the result is cached."""

    def __init__(self, backend, timestep_fs: float) -> None:
        if not timestep_fs > 0.0:
            raise ValueError("timestep_fs must be positive")
        self.backend = backend
        self.timestep_fs = float(timestep_fs)
        self.dt = self.timestep_fs * FS_TO_AU_TIME
        self.result = None

    def initialize(self, system: System):
        self.result = self.backend.compute(system.positions)
        return self.result

    def _accel(self, system: System) -> np.ndarray:
        return -np.asarray(self.result.gradient) / system.masses[:, None]

    def thermostat_energy(self) -> float:
        return 0.0

    def conserved_energy(self, system: System) -> float:
        if self.result is None:
            raise RuntimeError("integrator not initialised")
        return float(self.result.energy) + system.kinetic_energy() + self.thermostat_energy()

    def config(self) -> dict:
        return {"timestep_fs": self.timestep_fs}

    def state_dict(self) -> dict:
        return {"integrator": type(self).__name__, "config": self.config(), "result": _result_state(self.result)}

    def load_state_dict(self, state: dict) -> None:
        """Restore a saved state; the saved SCF density becomes the next SCF's guess when
        the backend starts each SCF from its previous density (aimd/integrators.py:125-163)."""
        if state.get("integrator") != type(self).__name__:
            raise ValueError(f"cannot load {state.get('integrator')!r} state into {type(self).__name__}")
        result = _result_from_state(state["result"])
        if (result is not None and result.density is not None
                and getattr(self.backend, "supports_density_guess", False)
                and getattr(self.backend, "reuse_density", True)):
            self.backend.set_density_guess(result.density)
        self.result = result


class VelocityVerlet(Integrator):
    """NVE velocity Verlet: v += dt/2 a; x += dt v; forces at x; v += dt/2 a."""

    def step(self, system: System):
        if self.result is None:
            self.initialize(system)
        system.velocities += 0.5 * self.dt * self._accel(system)
        system.positions += self.dt * system.velocities
        self.result = self._compute(system)
        system.velocities += 0.5 * self.dt * self._accel(system)
        return self.result

    def _compute(self, system: System):
        return self.backend.compute(system.positions)


# Niklasson, Steneteg, Odell, Bock, Challacombe, Tymczak, Holmstrom, Zheng and Weber,
# J. Chem. Phys. 130, 214109 (2009), Table I: K -> (kappa, alpha, c_0..c_K)
XL_TABLE = {
    3: (1.69, 150e-3, (-2, 3, 0, -1)),
    4: (1.75, 57e-3, (-3, 6, -2, -2, 1)),
    5: (1.82, 18e-3, (-6, 14, -8, -3, 4, -1)),
    6: (1.84, 5.5e-3, (-14, 36, -27, -2, 12, -6, 1)),
    7: (1.86, 1.6e-3, (-36, 99, -88, 11, 32, -25, 8, -1)),
    8: (1.88, 0.44e-3, (-99, 286, -286, 78, 78, -90, 42, -10, 1)),
    9: (1.89, 0.12e-3, (-286, 858, -936, 364, 168, -300, 184, -63, 12, -1)),
}


class AuxiliaryDensity:
    """
    History P(t), P(t-dt), ..., P(t-K dt) of the XL-BOMD auxiliary density and its
    dissipative Verlet step

        P(t+dt) = 2 P(t) - P(t-dt) + kappa (D(t) - P(t)) + alpha sum_k c_k P(t-k dt).

    The history is one stacked array, a torch tensor on the device when the
    densities handed in are device tensors, numpy otherwise; reset() starts it
    from rest at the first SCF density.
    """

    def __init__(self, k: int = 5) -> None:
        if k not in XL_TABLE:
            raise ValueError(f"XL-BOMD dissipation order K must be one of {sorted(XL_TABLE)}, got {k!r}")
        self.k = int(k)
        self.kappa, self.alpha, c = XL_TABLE[k]
        self.c = np.array(c, dtype=float)
        self.history = None

    @property
    def initialized(self) -> bool:
        return self.history is not None

    @property
    def current(self):
        if self.history is None:
            raise RuntimeError("auxiliary density not initialised")
        return self.history[0]

    def reset(self, density) -> None:
        if torch is not None and isinstance(density, torch.Tensor):
            d = density.detach().to(torch.float64)
            self.history = d.unsqueeze(0).repeat(self.k + 1, *([1] * d.dim())).contiguous()
        else:
            d = np.array(density, dtype=float)
            self.history = np.repeat(d[None], self.k + 1, axis=0)

    def propagate(self, density):
        """P(t+dt) from the SCF density D(t); shifts the history by one step."""
        h = self.history
        if h is None:
            raise RuntimeError("auxiliary density not initialised")
        if tuple(density.shape) != tuple(h.shape[1:]):
            raise ValueError(f"SCF density shape {tuple(density.shape)} != auxiliary density shape {tuple(h.shape[1:])}")
        if torch is not None and isinstance(h, torch.Tensor):
            d = torch.as_tensor(density, dtype=torch.float64, device=h.device)
            w = torch.as_tensor(self.c, dtype=torch.float64, device=h.device)
            damp = torch.tensordot(w, h, dims=([0], [0]))
            nxt = 2.0 * h[0] - h[1] + self.kappa * (d - h[0]) + self.alpha * damp
            self.history = torch.cat([nxt[None], h[:-1]])
            return nxt.clone()
        d = np.asarray(density, dtype=float)
        damp = np.tensordot(self.c, h, axes=(0, 0))
        nxt = 2.0 * h[0] - h[1] + self.kappa * (d - h[0]) + self.alpha * damp
        self.history = np.concatenate([nxt[None], h[:-1]])
        return nxt.copy()

    def state_dict(self) -> dict:
        h = self.history
        if torch is not None and isinstance(h, torch.Tensor):
            h = h.cpu().numpy()
        return {"k": self.k, "history": None if h is None else np.array(h)}

    def load_state_dict(self, state: dict, device=None) -> None:
        if int(state["k"]) != self.k:
            raise ValueError(f"checkpoint XL-BOMD order K={state['k']} != integrator's K={self.k}")
        h = state["history"]
        if h is None:
            self.history = None
            return
        h = np.array(h, dtype=float)
        if h.shape[0] != self.k + 1:
            raise ValueError(f"XL-BOMD history holds {h.shape[0]} densities, expected {self.k + 1}")
        self.history = torch.as_tensor(h, device=device) if device is not None else h


class XLBOMD(VelocityVerlet):
    """Velocity Verlet whose SCF guess at R(t+dt) is the propagated auxiliary density
    (aimd/integrators.py:321-416); the history starts from the first SCF density."""

    def __init__(self, backend, timestep_fs: float, k: int = 5) -> None:
        if not getattr(backend, "supports_density_guess", False):
            raise ValueError(f"XL-BOMD needs a backend that accepts SCF density guesses; "
                             f"{getattr(backend, 'name', type(backend).__name__)!r} does not")
        self.aux = AuxiliaryDensity(k)
        super().__init__(backend, timestep_fs)

    @property
    def k(self) -> int:
        return self.aux.k

    def _density(self, result):
        dev = getattr(self.backend, "device_density", None)
        if dev is not None and getattr(self.backend, "last_scf", None) is not None:
            return dev()
        if result.density is None:
            raise RuntimeError("the backend returned no SCF density, which XL-BOMD needs")
        return result.density

    def initialize(self, system: System):
        r = super().initialize(system)
        self.aux.reset(self._density(r))
        return r

    def _compute(self, system: System):
        if not self.aux.initialized:
            self.aux.reset(self._density(self.result))
        self.backend.set_density_guess(self.aux.propagate(self._density(self.result)))
        return self.backend.compute(system.positions)

    def config(self) -> dict:
        return {**super().config(), "k": self.k}

    def state_dict(self) -> dict:
        return {**super().state_dict(), "xl": self.aux.state_dict()}

    def load_state_dict(self, state: dict) -> None:
        aux = AuxiliaryDensity(self.k)
        dev = None
        if getattr(self.backend, "device_density", None) is not None and torch is not None:
            dev = self.backend.solver.device
        if state.get("integrator") == type(self).__name__:
            aux.load_state_dict(state["xl"], device=dev)
        super().load_state_dict(state)
        self.aux = aux


class LangevinBAOAB(Integrator):
    """
    Langevin dynamics, BAOAB splitting (Leimkuhler and Matthews 2013):
    B v += dt/2 a; A x += dt/2 v; O v = c1 v + c2 sqrt(kT/m) xi; A x += dt/2 v;
    forces at x; B v += dt/2 a, with c1 = exp(-gamma dt), c2 = sqrt(1 - c1^2).
    ``heat`` accumulates the kinetic energy the O step changes, so
    E_pot + E_kin - heat is conserved (aimd/integrators.py:419-491).
    """

    def __init__(self, backend, timestep_fs: float, temperature_k: float, friction_per_fs: float = 0.01,
                 rng=None) -> None:
        super().__init__(backend, timestep_fs)
        if not (math.isfinite(temperature_k) and temperature_k >= 0.0):
            raise ValueError("temperature_k must be finite and non-negative")
        if not (math.isfinite(friction_per_fs) and friction_per_fs >= 0.0):
            raise ValueError("friction_per_fs must be finite and non-negative")
        self.temperature_k = float(temperature_k)
        self.friction_per_fs = float(friction_per_fs)
        self.gamma = self.friction_per_fs / FS_TO_AU_TIME
        self.rng = np.random.default_rng(rng)
        self.c1 = math.exp(-self.gamma * self.dt)
        self.c2 = math.sqrt(1.0 - self.c1 ** 2)
        self.heat = 0.0

    def step(self, system: System):
        if self.result is None:
            self.initialize(system)
        half = 0.5 * self.dt
        system.velocities += half * self._accel(system)
        system.positions += half * system.velocities
        before = system.kinetic_energy()
        kick = np.sqrt(KB_AU * self.temperature_k / system.masses)[:, None] * self.rng.normal(
            size=system.velocities.shape)
        system.velocities = self.c1 * system.velocities + self.c2 * kick
        self.heat += system.kinetic_energy() - before
        system.com_removed = False          # random kicks thermalise all 3N momenta
        system.positions += half * system.velocities
        self.result = self.backend.compute(system.positions)
        system.velocities += half * self._accel(system)
        return self.result

    def thermostat_energy(self) -> float:
        return -self.heat

    def config(self) -> dict:
        return {"timestep_fs": self.timestep_fs, "temperature_k": self.temperature_k,
                "friction_per_fs": self.friction_per_fs}

    def state_dict(self) -> dict:
        return {**super().state_dict(), "heat": self.heat, "rng": self.rng.bit_generator.state}

    def load_state_dict(self, state: dict) -> None:
        if state.get("integrator") == type(self).__name__:
            heat = float(state["heat"])
            rng = np.random.default_rng()
            rng.bit_generator.state = _plain(state["rng"])
        super().load_state_dict(state)
        self.heat, self.rng = heat, rng


INTEGRATORS = {c.__name__: c for c in (VelocityVerlet, XLBOMD, LangevinBAOAB)}


def _plain(obj):
    """Arrays and numpy scalars of a decoded state back to Python ints (RNG states)."""
    if isinstance(obj, dict):
        return {k: _plain(v) for k, v in obj.items()}
    if isinstance(obj, np.ndarray) and obj.ndim == 0:
        return obj.item()
    return obj


class VelocityVerletNVE:
    """Array-level NVE driver (x, v held here) over :class:`VelocityVerlet`."""

    def __init__(self, backend, masses: np.ndarray, positions: np.ndarray, velocities: np.ndarray,
                 timestep_fs: float) -> None:
        self.system = System(["X"] * len(masses), np.array(positions, dtype=float),
                             np.array(velocities, dtype=float), np.asarray(masses, dtype=float))
        self.vv = VelocityVerlet(backend, timestep_fs)
        self.backend = backend
        self.dt = self.vv.dt

    @property
    def x(self) -> np.ndarray:
        return self.system.positions

    @property
    def v(self) -> np.ndarray:
        return self.system.velocities

    @property
    def m(self) -> np.ndarray:
        return self.system.masses

    @property
    def result(self):
        return self.vv.result

    def initialize(self):
        return self.vv.initialize(self.system)

    def step(self):
        return self.vv.step(self.system)

    def total_energy(self) -> float:
        return self.vv.conserved_energy(self.system)


# --------------------------------------------------------------- checkpoint
CHECKPOINT_FORMAT = "aimd-checkpoint"
CHECKPOINT_VERSION = 1
_META = "__metadata__"
_ARR = "__ndarray__"


def backend_fingerprint(backend) -> dict:
    """name, label, basis, nao and density_shape of a backend (aimd/checkpoint.py:123-149):
    a restart must stay on the same potential-energy surface."""
    out: dict[str, Any] = {"name": str(getattr(backend, "name", ""))}
    label = getattr(backend, "label", None)
    if isinstance(label, str):
        out["label"] = label
    basis = getattr(backend, "basis", None)
    basis = getattr(basis, "name", basis)
    if isinstance(basis, str):
        out["basis"] = basis.strip().lower()
    elif basis is not None:
        out["basis"] = "custom"
    nao = getattr(backend, "nao", None)
    if isinstance(nao, (int, np.integer)) and not isinstance(nao, bool):
        out["nao"] = int(nao)
    shape = getattr(backend, "density_shape", None)
    if shape is not None:
        out["density_shape"] = [int(n) for n in shape]
    return out


def _pack(obj, arrays: dict):
    if isinstance(obj, np.ndarray):
        key = f"a{len(arrays):04d}"
        arrays[key] = obj
        return {_ARR: key}
    if isinstance(obj, dict):
        return {str(k): _pack(v, arrays) for k, v in obj.items()}
    if isinstance(obj, (list, tuple)):
        return [_pack(v, arrays) for v in obj]
    if isinstance(obj, np.generic):
        return obj.item()
    if obj is None or isinstance(obj, (bool, int, float, str)):
        return obj
    raise TypeError(f"cannot checkpoint object of type {type(obj).__name__}")


def _unpack(obj, arrays):
    if isinstance(obj, dict):
        if set(obj) == {_ARR}:
            return np.array(arrays[obj[_ARR]])
        return {k: _unpack(v, arrays) for k, v in obj.items()}
    if isinstance(obj, list):
        return [_unpack(v, arrays) for v in obj]
    return obj


def _checkpointable_info(state: dict) -> dict:
    res = state.get("result")
    if not res or not res.get("info"):
        return state
    keep = {}
    for k, v in res["info"].items():
        try:
            _pack(v, {})
            keep[k] = v
        except TypeError:
            pass
    return {**state, "result": {**res, "info": keep}}


@dataclass
class Checkpoint:
    system: System
    step: int
    time_fs: float
    integrator_state: dict
    backend: str = ""
    backend_info: dict = field(default_factory=dict)
    run_info: dict = field(default_factory=dict)

    def check_backend(self, backend) -> None:
        mine = backend_fingerprint(backend)
        diff = [f"{k}: {self.backend_info.get(k)!r} -> {mine.get(k)!r}"
                for k in sorted(set(self.backend_info) | set(mine)) if self.backend_info.get(k) != mine.get(k)]
        if self.backend_info and diff:
            raise ValueError("the backend differs from the one that wrote the checkpoint (" + "; ".join(diff) + ")")

    def restore(self, system: System, integrator: Integrator) -> None:
        if list(system.symbols) != list(self.system.symbols):
            raise ValueError(f"checkpoint atoms {self.system.symbols} != system atoms {system.symbols}")
        integrator.load_state_dict(self.integrator_state)
        system.positions = self.system.positions.copy()
        system.velocities = self.system.velocities.copy()
        system.masses = self.system.masses.copy()
        system.com_removed = self.system.com_removed


def save_checkpoint(path, system: System, integrator: Integrator, step: int, time_fs: float | None = None,
                    run_info: dict | None = None) -> Path:
    """Write system + integrator state after ``step`` steps; atomic (temporary file + rename)."""
    path = Path(path)
    meta = {
        "format": CHECKPOINT_FORMAT, "version": CHECKPOINT_VERSION, "step": int(step),
        "time_fs": float(step * integrator.timestep_fs if time_fs is None else time_fs),
        "backend": str(getattr(integrator.backend, "name", "")),
        "backend_info": backend_fingerprint(integrator.backend),
        "run": dict(run_info or {}),
        "system": {"symbols": list(system.symbols), "positions": np.array(system.positions, dtype=float),
                   "velocities": np.array(system.velocities, dtype=float),
                   "masses": np.array(system.masses, dtype=float), "charge": 0, "multiplicity": 1,
                   "com_removed": bool(system.com_removed), "rotation_removed": False, "rotational_dof": None},
        "integrator": _checkpointable_info(integrator.state_dict()),
    }
    arrays: dict[str, np.ndarray] = {}
    packed = _pack(meta, arrays)
    arrays[_META] = np.array(json.dumps(packed))
    tmp = path.with_name(path.name + ".tmp")
    with tmp.open("wb") as fh:
        np.savez(fh, **arrays)
        fh.flush()
        os.fsync(fh.fileno())
    os.replace(tmp, path)
    return path


def load_checkpoint(path) -> Checkpoint:
    """Read a checkpoint (numpy.load with allow_pickle=False: nothing in it executes)."""
    try:
        with np.load(Path(path), allow_pickle=False) as npz:
            if _META not in npz.files:
                raise ValueError(f"{path} is not an aimd checkpoint")
            meta = json.loads(str(npz[_META]))
            if not isinstance(meta, dict) or meta.get("format") != CHECKPOINT_FORMAT:
                raise ValueError(f"{path} is not an aimd checkpoint")
            if meta.get("version") != CHECKPOINT_VERSION:
                raise ValueError(f"unsupported checkpoint version {meta.get('version')}")
            data = _unpack(meta, {k: npz[k] for k in npz.files if k != _META})
    except (zipfile.BadZipFile, EOFError, json.JSONDecodeError, KeyError, TypeError) as e:
        raise ValueError(f"{path} is not a valid aimd checkpoint ({type(e).__name__}: {e})") from e
    s = data["system"]
    system = System(list(s["symbols"]), np.array(s["positions"], dtype=float),
                    np.array(s["velocities"], dtype=float), np.array(s["masses"], dtype=float),
                    bool(s["com_removed"]))
    return Checkpoint(system, int(data["step"]), float(data["time_fs"]), data["integrator"],
                      data.get("backend", ""), data.get("backend_info") or {}, data.get("run") or {})


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
