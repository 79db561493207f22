"""
Device-side integral engine: one plan per basis, one geometry at a time.

Wraps the C ABI (include/aimdhf.h) for the Python host layer. Matrices live
in PyTorch tensors on the current HIP device (float64, row-major); every
call is enqueued on torch's current stream.
"""

from __future__ import annotations

import ctypes

import numpy as np
import torch

from aimd_amd import _lib
from aimd_amd.qc.basis import BasisSet

DEFAULT_SCHWARZ = 1e-12          # aimd/qc/integrals.py:114
DEFAULT_SCHWARZ_GRAD = 1e-14     # aimd/qc/integrals.py:119


def _ptr(a) -> ctypes.c_void_p:
    if isinstance(a, torch.Tensor):
        return ctypes.c_void_p(a.data_ptr())
    return a.ctypes.data_as(ctypes.c_void_p)


def _stream() -> ctypes.c_void_p:
    return ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)


def require_device() -> torch.device:
    if not torch.cuda.is_available():
        raise RuntimeError("the aimd_amd engine needs a ROCm GPU (torch.cuda.is_available() is False); "
                           "there is no CPU fallback")
    return torch.device("cuda", torch.cuda.current_device())


class Engine:
    """Integral plan of one basis on the current device."""

    def __init__(self, basis: BasisSet) -> None:
        self.device = require_device()
        L = _lib.lib()
        self.basis = basis
        self.nao = basis.nao
        self.natm = basis.natm
        self._keep = [np.ascontiguousarray(a) for a in (
            basis.shell_atom.astype(np.int64), basis.shell_l.astype(np.int64),
            basis.shell_prim.astype(np.int64), basis.shell_block.astype(np.int64),
            basis.prim_exp.astype(np.float64), basis.prim_coef.astype(np.float64),
            basis.nuclear_charges.astype(np.float64))]
        h = ctypes.c_void_p()
        _lib.check(L.aimdhf_plan_create(self.natm, basis.nshell, *[_ptr(a) for a in self._keep],
                                        ctypes.byref(h)))
        self._h = h
        self.positions: np.ndarray | None = None
        self.threshold: float | None = None

    def close(self) -> None:
        if getattr(self, "_h", None):
            _lib.lib().aimdhf_plan_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # ------------------------------------------------------------ geometry
    def set_geometry(self, positions, schwarz_threshold: float = DEFAULT_SCHWARZ) -> None:
        x = np.ascontiguousarray(np.asarray(positions, dtype=np.float64).reshape(self.natm, 3))
        _lib.check(_lib.lib().aimdhf_set_geometry(self._h, _ptr(x), float(schwarz_threshold), _stream()))
        self.positions = x.copy()
        self.threshold = float(schwarz_threshold)

    def schwarz_bounds(self) -> np.ndarray:
        nb = self.basis.nblock
        out = np.zeros(nb * (nb + 1) // 2)
        _lib.check(_lib.lib().aimdhf_schwarz_bounds(self._h, _ptr(out), out.size))
        return out

    def quartet_count(self) -> int:
        n = ctypes.c_int64()
        _lib.check(_lib.lib().aimdhf_quartet_count(self._h, ctypes.byref(n)))
        return int(n.value)

    def quartets(self) -> np.ndarray:
        """Surviving unique shell quartets as (P, Q) shell-pair indices, P >= Q."""
        n = self.quartet_count()
        out = np.zeros((max(n, 1), 2), np.int64)
        cnt = ctypes.c_int64()
        _lib.check(_lib.lib().aimdhf_quartets(self._h, _ptr(out), n, ctypes.byref(cnt)))
        return out[:n]

    # ---------------------------------------------------------------- J / K
    JK_MODES = {"auto": 0, "direct": 1, "store": 2}

    def set_jk_mode(self, mode: str = "auto", max_store_bytes: int = 0) -> None:
        """J/K mode from the next set_geometry on: "direct" recomputes the integrals every
        build, "store" evaluates them once per geometry into device memory, "auto" stores
        when the estimated size is within max_store_bytes (0: half the free device memory)."""
        if mode not in self.JK_MODES:
            raise ValueError(f"jk mode must be one of {sorted(self.JK_MODES)}, got {mode!r}")
        _lib.check(_lib.lib().aimdhf_set_jk_mode(self._h, self.JK_MODES[mode], int(max_store_bytes)))
        self.jk_mode = mode

    def set_store_units(self, items: int) -> None:
        """Lane items per store-mode digestion work unit (from the next set_geometry)."""
        _lib.check(_lib.lib().aimdhf_set_store_units(self._h, int(items)))
        self.store_units = int(items)

    PRECISIONS = {"fp64": 0, "mixed": 1}

    def set_precision(self, precision: str = "fp64", jk_switch: float = 0.0, grad_switch: float = 0.0,
                      delta_switch: float = 0.0) -> None:
        """"fp64", or "mixed": FP32 for quartets whose Schwarz bound (J/K store, from the next
        set_geometry) or Schwarz-Gamma bound (derivative integrals) is below the switch, and
        for every stored integral in density-difference builds with max |dD| <= delta_switch."""
        if precision not in self.PRECISIONS:
            raise ValueError(f"precision must be one of {sorted(self.PRECISIONS)}, got {precision!r}")
        _lib.check(_lib.lib().aimdhf_set_precision(self._h, self.PRECISIONS[precision], float(jk_switch),
                                                   float(grad_switch), float(delta_switch)))
        self.precision = precision

    def set_streams(self, concurrent: bool = True) -> None:
        """Spread independent launches over extra library streams (default) or keep every
        launch on torch's current stream (per-kernel timings then exclude overlap)."""
        _lib.check(_lib.lib().aimdhf_set_streams(self._h, int(bool(concurrent))))

    def jk_info(self) -> dict:
        """Store state of the current geometry: stored, bytes, K and J gather-list entries and
        the bytes of integral data one J/K build streams (gather_bytes)."""
        st = ctypes.c_int()
        b, ke, je, gd = ctypes.c_int64(), ctypes.c_int64(), ctypes.c_int64(), ctypes.c_int64()
        _lib.check(_lib.lib().aimdhf_jk_info(self._h, ctypes.byref(st), ctypes.byref(b), ctypes.byref(ke),
                                             ctypes.byref(je), ctypes.byref(gd)))
        return {"stored": bool(st.value), "store_bytes": int(b.value), "k_entries": int(ke.value),
                "j_entries": int(je.value), "gather_bytes": int(gd.value)}

    def jk_type_info(self) -> dict:
        """Store traffic per exchange output type (9), Coulomb pair class (6) and the unique
        integral bytes per quartet class (21) of the current geometry (zeros in direct mode)."""
        kb, ke = np.zeros(9, np.int64), np.zeros(9, np.int64)
        jb, je = np.zeros(6, np.int64), np.zeros(6, np.int64)
        ub = np.zeros(21, np.int64)
        _lib.check(_lib.lib().aimdhf_jk_type_info(self._h, _ptr(kb), _ptr(ke), _ptr(jb), _ptr(je), _ptr(ub)))
        return {"k_bytes": kb, "k_entries": ke, "j_bytes": jb, "j_entries": je, "unique_bytes": ub}

    def jk(self, dens: torch.Tensor) -> tuple[torch.Tensor, torch.Tensor]:
        """J, K of (nset, nao, nao) (or (nao, nao)) densities; symmetric part of each D is used."""
        single = dens.dim() == 2
        D = dens.reshape(-1, self.nao, self.nao).to(device=self.device, dtype=torch.float64).contiguous()
        J = torch.empty_like(D)
        K = torch.empty_like(D)
        _lib.check(_lib.lib().aimdhf_jk(self._h, D.shape[0], _ptr(D), _ptr(J), _ptr(K), _stream()))
        return (J[0], K[0]) if single else (J, K)

    def jk_delta(self, ddens: torch.Tensor, threshold: float) -> tuple[torch.Tensor, torch.Tensor]:
        """J, K of density differences for an incremental Fock build.

        In store mode, work units whose largest Schwarz product times max |dD|
        is below ``threshold`` are skipped, and in mixed precision the integrals
        are digested in FP32. In direct mode this is :meth:`jk`.
        """
        single = ddens.dim() == 2
        D = ddens.reshape(-1, self.nao, self.nao).to(device=self.device, dtype=torch.float64).contiguous()
        J = torch.empty_like(D)
        K = torch.empty_like(D)
        _lib.check(_lib.lib().aimdhf_jk_delta(self._h, D.shape[0], _ptr(D), _ptr(J), _ptr(K), float(threshold),
                                              _stream()))
        return (J[0], K[0]) if single else (J, K)

    # ------------------------------------------------------- one-electron
    def one_electron(self, S: torch.Tensor, T: torch.Tensor, V: torch.Tensor, M: torch.Tensor) -> None:
        """Fill S, T, V (nao, nao) and the dipole integrals M (3, nao, nao) of the current geometry."""
        for t, shp in ((S, (self.nao, self.nao)), (T, (self.nao, self.nao)), (V, (self.nao, self.nao)),
                       (M, (3, self.nao, self.nao))):
            if tuple(t.shape) != shp or t.dtype != torch.float64 or not t.is_contiguous() or t.device != self.device:
                raise ValueError(f"output tensor must be a contiguous float64 {shp} tensor on {self.device}")
        _lib.check(_lib.lib().aimdhf_one_electron(self._h, _ptr(S), _ptr(T), _ptr(V), _ptr(M), _stream()))

    # ----------------------------------------------------------- gradients
    def grad_one(self, D: torch.Tensor, W: torch.Tensor) -> torch.Tensor:
        """d/dR [tr(D h) - tr(W S)] as a (natm, 3) device tensor."""
        Dd = D.to(device=self.device, dtype=torch.float64).contiguous()
        Wd = W.to(device=self.device, dtype=torch.float64).contiguous()
        g = torch.empty((self.natm, 3), dtype=torch.float64, device=self.device)
        _lib.check(_lib.lib().aimdhf_grad_one(self._h, _ptr(Dd), _ptr(Wd), _ptr(g), _stream()))
        return g

    def grad_one_noalloc(self, D: torch.Tensor, W: torch.Tensor, out: torch.Tensor) -> torch.Tensor:
        """grad_one into the caller's (natm, 3) device tensor ``out`` through the entry point
        that allocates no device memory (fixed library scratch; nao <= 1024)."""
        if tuple(out.shape) != (self.natm, 3) or out.dtype != torch.float64 or not out.is_contiguous():
            raise ValueError(f"out must be a contiguous float64 ({self.natm}, 3) tensor")
        _lib.check(_lib.lib().aimdhf_grad_one_noalloc(self._h, _ptr(D), _ptr(W), _ptr(out), _stream()))
        return out

    def grad_one_terms(self, D: torch.Tensor, W: torch.Tensor) -> tuple[torch.Tensor, torch.Tensor, torch.Tensor]:
        """(d/dR tr(D T), d/dR tr(D V), -d/dR tr(W S)), each a (natm, 3) device tensor."""
        Dd = D.to(device=self.device, dtype=torch.float64).contiguous()
        Wd = W.to(device=self.device, dtype=torch.float64).contiguous()
        g = torch.empty((3, self.natm, 3), dtype=torch.float64, device=self.device)
        _lib.check(_lib.lib().aimdhf_grad_one_terms(self._h, _ptr(Dd), _ptr(Wd), _ptr(g[0]), _ptr(g[1]), _ptr(g[2]),
                                                    _stream()))
        return g[0], g[1], g[2]

    def grad_two(self, Dc: torch.Tensor, Dx: torch.Tensor, kfac: float,
                 threshold: float = DEFAULT_SCHWARZ_GRAD) -> torch.Tensor:
        """d/dR of the two-electron energy at fixed densities, (natm, 3) device tensor."""
        Dc = Dc.to(device=self.device, dtype=torch.float64).contiguous()
        Dx = Dx.reshape(-1, self.nao, self.nao).to(device=self.device, dtype=torch.float64).contiguous()
        g = torch.empty((self.natm, 3), dtype=torch.float64, device=self.device)
        _lib.check(_lib.lib().aimdhf_grad_two(self._h, _ptr(Dc), Dx.shape[0], _ptr(Dx), float(kfac),
                                              float(threshold), _ptr(g), _stream()))
        return g

    # ------------------------------------------------------------ profiling
    PAIR_CLASS_NAMES = ["ss", "ps", "pp", "ds", "dp", "dd"]
    KIND_NAMES = ([f"jengine_bra_{c}" for c in PAIR_CLASS_NAMES]
                  + [f"exchange_{'spd'[t // 3]}{'spd'[t % 3]}" for t in range(9)]
                  + ["gather_copies"] + [f"unused_{k}" for k in range(16, 21)]
                  + [f"grad_class_{k}" for k in range(21)]
                  + ["geometry", "one_electron", "grad_one", "jk_prologue_epilogue", "jk_hermite_density",
                     "eri_store", "gather_lists"]
                  + [f"jstore_{c}" for c in PAIR_CLASS_NAMES]
                  + [f"kstore_{'spd'[t // 3]}{'spd'[t % 3]}" for t in range(9)]
                  + [f"grad_screen_{k}" for k in range(21)]
                  + ["host_energy_list", "host_owner_lists", "host_store_build", "host_gradient_list"]
                  + ["host_store_counts", "host_store_bands", "host_store_fill", "host_store_units"]
                  + ["kstore_reduce", "jstore_reduce", "generic_jk", "generic_grad_two", "unused_97"])
    NKIND = 98
    FOCK_KINDS = list(range(0, 15)) + [45, 46] + list(range(49, 64)) + [93, 94, 95]   # per-iteration J/K work
    STORE_KINDS = [15, 47, 48]                                          # per-geometry store build
    GRAD_KINDS = list(range(21, 42)) + list(range(64, 85)) + [96]       # derivative ERIs

    def profile(self, on: bool = True) -> None:
        """Record HIP events around every engine launch (on torch's current stream)."""
        _lib.check(_lib.lib().aimdhf_profile_enable(self._h, int(bool(on))))

    def profile_read(self, reset: bool = False) -> tuple[np.ndarray, np.ndarray]:
        """(milliseconds, launch counts) per kind, NKIND kinds (see KIND_NAMES)."""
        ms = np.zeros(self.NKIND)
        n = np.zeros(self.NKIND, np.int64)
        _lib.check(_lib.lib().aimdhf_profile_read(self._h, _ptr(ms), _ptr(n), self.NKIND, int(bool(reset))))
        return ms, n

    def profile_replay(self, kind: int, reps: int = 20) -> float:
        """Average milliseconds of `reps` back-to-back launches of a store-mode digestion
        kernel (kind 49..63) with the arguments of the last J/K build."""
        out = ctypes.c_double()
        _lib.check(_lib.lib().aimdhf_profile_replay(self._h, int(kind), int(reps), _stream(), ctypes.byref(out)))
        return float(out.value)

    def class_counts(self) -> tuple[np.ndarray, np.ndarray]:
        """Per-class sizes of the energy and (last) gradient quartet lists."""
        e = np.zeros(21, np.int64)
        g = np.zeros(21, np.int64)
        _lib.check(_lib.lib().aimdhf_class_counts(self._h, _ptr(e), _ptr(g)))
        return e, g

    def grad_quartets(self) -> np.ndarray:
        """Quartets differentiated by the last grad_two (after its Gamma screening), (n, 2), P >= Q."""
        cnt = ctypes.c_int64()
        _lib.check(_lib.lib().aimdhf_grad_quartets(self._h, None, 0, ctypes.byref(cnt)))
        out = np.zeros((max(cnt.value, 1), 2), np.int64)
        _lib.check(_lib.lib().aimdhf_grad_quartets(self._h, _ptr(out), cnt.value, ctypes.byref(cnt)))
        return out[:cnt.value]

    def pair_npp(self) -> np.ndarray:
        npair = self.basis.nshell * (self.basis.nshell + 1) // 2
        out = np.zeros(npair, np.int32)
        _lib.check(_lib.lib().aimdhf_pair_npp(self._h, _ptr(out), npair))
        return out
