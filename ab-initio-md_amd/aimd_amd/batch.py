"""
Batched replicas: many independent molecules of uneven size stepped together
on one GPU through one batched engine plan (``aimdhf_batch_create``).

Per geometry of the whole batch:

* one set-up (pairs, Schwarz bounds, screened quartets of every replica) and
  one integral-evaluation launch per quartet class for all replicas, expanded
  into one dense two-electron tensor per replica;
* an SCF whose iteration is one captured HIP graph for all replicas: J/K of
  every replica in one launch, Fock matrices, energies and commutators as
  batched products, Pulay DIIS with its small systems solved on the device
  (one launch), and the aufbau densities of every replica and spin by
  purification in one launch; one host read per iteration checks convergence;
* forces of every replica from one one-electron-gradient launch and one
  derivative-ERI launch per class.

Every replica is computed in the unrestricted form (alpha and beta densities):
a closed-shell replica starts with equal spin densities and keeps them equal,
so its energy and forces are those of RHF; open-shell replicas (the methyl
radical) are UHF. Matrices use the padded layout of the engine: replica r's
ld x ld block, ld the largest replica's AO count, padding rows and columns
zero (the padded overlap diagonal is one, and the padded orbitals are shifted
far above the Fermi level so they stay empty).

The single-molecule semantics each replica reproduces are those of the ``hf``
backend's compute (aimd/backends/hf.py:293-363); :class:`BatchLangevinBAOAB`
steps all replicas with the reference's BAOAB splitting
(aimd/integrators.py:450-471) on device tensors.
"""

from __future__ import annotations

import ctypes
import math
from typing import Sequence

import numpy as np
import torch

from aimd_amd import _lib
from aimd_amd.md import masses_au
from aimd_amd.qc import engine as qengine
from aimd_amd.qc.basis import BasisSet, build_basis
from aimd_amd.qc.scf import electron_counts, nuclear_repulsion_energy, nuclear_repulsion_gradient, sad_density
from aimd_amd.units import FS_TO_AU_TIME, KB_AU

PAD_LEVEL = 50.0        # orbital energy (Eh) given to the padded AOs: far above any LUMO here


def _ptr(t: torch.Tensor) -> ctypes.c_void_p:
    return ctypes.c_void_p(t.data_ptr())


def _stream() -> ctypes.c_void_p:
    return ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)


class ReplicaBatch:
    """
    RHF/UHF energies and forces of many molecules at once.

    ``molecules``: (symbols, charge, multiplicity) per replica. Positions are
    handed to :meth:`compute` concatenated replica by replica ((natm_total, 3),
    bohr). Options as the ``hf`` backend's: SCF convergence on the energy
    change and the largest commutator element, Schwarz threshold, gradient
    screening; ``precision`` "fp64" or "mixed" (FP32 for quartets below the
    switch thresholds in the integral evaluation and the derivative
    integrals, as the single-molecule engine). ``jk_lanes`` (16, 32 or 64)
    fixes how many lanes digest one J/K element; None lets :meth:`autotune`
    time the choices on the first geometry. The choice sets the summation
    order, so runs that must agree bit for bit across processes fix it.
    """

    def __init__(self, molecules: Sequence[tuple], basis: str = "6-31g**", conv_tol: float = 1e-10,
                 conv_tol_grad: float = 1e-7, max_cycles: int = 100, diis_space: int = 8,
                 schwarz_threshold: float = qengine.DEFAULT_SCHWARZ,
                 gradient_screening: float = qengine.DEFAULT_SCHWARZ_GRAD, precision: str = "fp64",
                 jk_fp32_switch: float = 1e-5, grad_fp32_switch: float = 1e-5, purify_steps: int = 60,
                 lindep_threshold: float = 1e-7, capture: bool = True, jk_lanes: int | None = None) -> None:
        if not molecules:
            raise ValueError("a batch needs at least one molecule")
        if not 2 <= diis_space <= 12:
            raise ValueError("diis_space must be in 2..12")
        self.device = qengine.require_device()
        self.conv_tol, self.conv_tol_grad = float(conv_tol), float(conv_tol_grad)
        self.max_cycles, self.diis_space, self.purify_steps = int(max_cycles), int(diis_space), int(purify_steps)
        self.schwarz, self.gradient_screening = float(schwarz_threshold), float(gradient_screening)
        self.lindep = float(lindep_threshold)
        self.capture = bool(capture)            # replay the SCF iteration as one captured graph
        self.molecules = [(list(s), int(c), int(m)) for s, c, m in molecules]
        self.bases: list[BasisSet] = []
        cache: dict[tuple, BasisSet] = {}
        for sym, _, _ in self.molecules:
            key = tuple(sym)
            if key not in cache:
                cache[key] = build_basis(sym, np.zeros((len(sym), 3)), basis)
            self.bases.append(cache[key])
        self.nrep = len(self.molecules)
        self.natm_rep = np.array([b.natm for b in self.bases])
        self.rep_atom = np.concatenate([[0], np.cumsum(self.natm_rep)]).astype(np.int64)
        self.natm = int(self.rep_atom[-1])
        self.symbols = [s for sym, _, _ in self.molecules for s in sym]
        self.charges = np.concatenate([b.nuclear_charges for b in self.bases])
        self.masses = masses_au(self.symbols)
        occ = []
        for b, (sym, charge, mult) in zip(self.bases, self.molecules):
            occ.append(electron_counts(int(b.atomic_numbers.sum()), charge, mult, "uhf"))
        self.nocc = np.array(occ, dtype=float)                      # (nrep, 2): alpha, beta
        self.unrestricted = np.array([m != 1 for _, _, m in self.molecules])
        self._create_plan(precision, jk_fp32_switch, grad_fp32_switch)
        self.jk_lanes = None
        self.jk_times: dict[int, float] = {}
        if jk_lanes is not None:
            _lib.check(_lib.lib().aimdhf_batch_set_jk_lanes(self._h, int(jk_lanes)))
            self.jk_lanes = int(jk_lanes)
        self._alloc()
        self._graph = None
        self.Ds = None                                              # (nrep, 2, ld, ld) last SCF densities

    # ------------------------------------------------------------- plan
    def _create_plan(self, precision, jk_switch, grad_switch) -> None:
        arrs = {k: [] for k in ("atom", "l", "prim", "block", "exp", "coef")}
        rep_shell = [0]
        aoff = poff = boff = 0
        for b in self.bases:
            arrs["atom"].append(b.shell_atom + aoff)
            arrs["l"].append(b.shell_l)
            arrs["prim"].append(b.shell_prim[:-1] + poff)
            arrs["block"].append(b.shell_block + boff)
            arrs["exp"].append(b.prim_exp)
            arrs["coef"].append(b.prim_coef)
            aoff += b.natm
            poff += b.nprim
            boff += b.nblock
            rep_shell.append(rep_shell[-1] + b.nshell)
        self._keep = [np.ascontiguousarray(a) for a in (
            self.rep_atom, np.array(rep_shell, np.int64),
            np.concatenate(arrs["atom"]).astype(np.int64), np.concatenate(arrs["l"]).astype(np.int64),
            np.concatenate(arrs["prim"] + [np.array([poff])]).astype(np.int64),
            np.concatenate(arrs["block"]).astype(np.int64), np.concatenate(arrs["exp"]).astype(np.float64),
            np.concatenate(arrs["coef"]).astype(np.float64), self.charges.astype(np.float64))]
        L = _lib.lib()
        h = ctypes.c_void_p()
        k = self._keep
        _lib.check(L.aimdhf_batch_create(self.nrep, k[0].ctypes.data_as(ctypes.c_void_p),
                                         k[1].ctypes.data_as(ctypes.c_void_p), self.natm, int(rep_shell[-1]),
                                         *[a.ctypes.data_as(ctypes.c_void_p) for a in k[2:]], ctypes.byref(h)))
        self._h = h
        if precision not in qengine.Engine.PRECISIONS:
            raise ValueError(f"precision must be one of {sorted(qengine.Engine.PRECISIONS)}, got {precision!r}")
        self.precision = precision
        _lib.check(L.aimdhf_set_precision(h, qengine.Engine.PRECISIONS[precision], float(jk_switch),
                                          float(grad_switch), 0.0))
        nrep, ld, tb = ctypes.c_int(), ctypes.c_int(), ctypes.c_int64()
        _lib.check(L.aimdhf_batch_layout(h, ctypes.byref(nrep), ctypes.byref(ld), ctypes.byref(tb)))
        self.ld, self.tensor_bytes = int(ld.value), int(tb.value)
        if self.ld > 64:
            raise ValueError(f"batched replicas need at most 64 AOs each (largest has {self.ld})")
        self.rep_nao = np.array([b.nao for b in self.bases])

    def close(self) -> None:
        if getattr(self, "_h", None):
            _lib.lib().aimdhf_plan_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # ---------------------------------------------------------- buffers
    def _alloc(self) -> None:
        R, n, M, dev = self.nrep, self.ld, self.diis_space, self.device
        f64 = dict(dtype=torch.float64, device=dev)
        self.S = torch.zeros((R, n, n), **f64)
        self.T = torch.zeros((R, n, n), **f64)
        self.V = torch.zeros((R, n, n), **f64)
        self.Mdip = torch.zeros((3, R, n, n), **f64)
        self.H = torch.zeros((R, n, n), **f64)
        self.X = torch.zeros((R, n, n), **f64)
        self.D = torch.zeros((R, 2, n, n), **f64)
        self.J = torch.zeros((R, 2, n, n), **f64)
        self.K = torch.zeros((R, 2, n, n), **f64)
        self.F = torch.zeros((R, 2, n, n), **f64)
        self.Fbuf = torch.zeros((R, M, 2, n, n), **f64)
        self.Ebuf = torch.zeros((R, M, 2, n, n), **f64)
        self.Pbuf = torch.zeros((R * 2, n, n), **f64)
        self.perr = torch.zeros(R * 2, **f64)
        self.G = torch.zeros((R, M, M), **f64)
        self.rhs = torch.zeros((R, M), **f64)
        self.coef = torch.zeros((R, M), **f64)
        self.energy_el = torch.zeros(R, **f64)
        self.e_prev = torch.full((R,), float("nan"), **f64)
        self.err = torch.zeros(R, **f64)
        self.frozen = torch.zeros(R, dtype=torch.bool, device=dev)
        self.slot = torch.zeros(1, dtype=torch.long, device=dev)
        self.count = torch.zeros(1, dtype=torch.long, device=dev)
        self.nocc_t = torch.tensor(self.nocc.reshape(-1), **f64)
        idx = torch.arange(n, device=dev)
        real = idx[None, :] < torch.tensor(self.rep_nao, device=dev)[:, None]          # (R, n)
        self.real = real
        self.pad_diag = torch.diag_embed((~real).to(torch.float64))                     # (R, n, n)
        self.eye_m = torch.eye(M, **f64)
        self.slots = torch.arange(M, device=dev)

    # ------------------------------------------------------- integrals
    def _set_geometry(self, x: np.ndarray) -> None:
        L = _lib.lib()
        xc = np.ascontiguousarray(np.asarray(x, dtype=np.float64).reshape(self.natm, 3))
        if not np.all(np.isfinite(xc)):
            raise ValueError("atomic positions must be finite (got NaN or inf)")
        _lib.check(L.aimdhf_set_geometry(self._h, xc.ctypes.data_as(ctypes.c_void_p), self.schwarz, _stream()))
        _lib.check(L.aimdhf_one_electron(self._h, _ptr(self.S), _ptr(self.T), _ptr(self.V), _ptr(self.Mdip),
                                         _stream()))
        self.S.add_(self.pad_diag)
        torch.add(self.T, self.V, out=self.H)
        self.X.copy_(self._inv_sqrt(self.S))
        self.positions = xc
        self.e_nuc = np.array([nuclear_repulsion_energy(self.charges[a:b], xc[a:b])
                               for a, b in zip(self.rep_atom[:-1], self.rep_atom[1:])])

    def _inv_sqrt(self, S: torch.Tensor, iters: int = 30) -> torch.Tensor:
        """S^(-1/2) of every replica by the coupled Newton-Schulz iteration (batched products);
        the residual |X S X - 1| is checked once per geometry on the host."""
        eye = torch.eye(self.ld, dtype=S.dtype, device=S.device).expand_as(S)
        c = S.abs().sum(dim=2).amax(dim=1)[:, None, None]
        Y = S / c
        Z = eye.clone()
        for _ in range(iters):
            Tm = 0.5 * (3.0 * eye - Z @ Y)
            Y = Y @ Tm
            Z = Tm @ Z
        Xs = Z / torch.sqrt(c)
        Xs = 0.5 * (Xs + Xs.transpose(1, 2))
        res = float((Xs @ S @ Xs - eye).abs().amax())
        if not res < 1e-9:
            s, U = torch.linalg.eigh(S)
            if float(s.amin()) < self.lindep:
                raise ValueError(f"overlap eigenvalue {float(s.amin()):.3e} below the linear-dependence threshold")
            Xs = (U * s.rsqrt()[:, None, :]) @ U.transpose(1, 2)
        return Xs

    # ------------------------------------------------------------ SCF
    def _jk(self, D: torch.Tensor) -> None:
        _lib.check(_lib.lib().aimdhf_jk(self._h, 2, _ptr(D), _ptr(self.J), _ptr(self.K), _stream()))

    def autotune(self, reps: int = 5) -> dict[int, float]:
        """Time the batched J/K digestion at 64, 32 and 16 lanes per output element on the
        current geometry's integral tensors (reps builds each, HIP events on the current
        stream), keep the fastest, and drop the captured SCF graph so the next iteration
        records the new launch. Returns {lanes: ms per build}."""
        L = _lib.lib()
        D = self.Ds if self.Ds is not None else self._guess()
        times = {}
        for lanes in (64, 32, 16):
            _lib.check(L.aimdhf_batch_set_jk_lanes(self._h, lanes))
            self._jk(D)
            a, b = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
            a.record()
            for _ in range(reps):
                self._jk(D)
            b.record()
            b.synchronize()
            times[lanes] = a.elapsed_time(b) / reps
        best = min(times, key=times.get)
        _lib.check(L.aimdhf_batch_set_jk_lanes(self._h, best))
        self.jk_lanes, self.jk_times = best, times
        self._graph = None
        return times

    def _iteration(self) -> None:
        """One SCF iteration of every replica on static buffers (captured as a graph)."""
        L = _lib.lib()
        R, n, M = self.nrep, self.ld, self.diis_space
        D, X, S, H = self.D, self.X, self.S, self.H
        self._jk(D)
        Jt = self.J[:, 0] + self.J[:, 1]
        F = H[:, None] + Jt[:, None] - self.K
        F = 0.5 * (F + F.transpose(2, 3))
        self.F.copy_(F)
        self.energy_el.copy_(0.5 * (D * (H[:, None] + F)).sum(dim=(1, 2, 3)))
        Xe = X[:, None]
        FDS = F @ D @ S[:, None]
        err = Xe.transpose(2, 3) @ (FDS - FDS.transpose(2, 3)) @ Xe
        self.err.copy_(err.abs().amax(dim=(1, 2, 3)))
        Fo = Xe.transpose(2, 3) @ F @ Xe
        # DIIS: push into the ring slot, Gram matrix of the stored errors, anchored system
        self.Fbuf.index_copy_(1, self.slot, Fo[:, None])
        self.Ebuf.index_copy_(1, self.slot, err[:, None])
        Ev = self.Ebuf.view(R, M, -1)
        G = Ev @ Ev.transpose(1, 2)                                                     # (R, M, M)
        a = self.slot.view(1, 1, 1).expand(R, M, 1)
        Gia = torch.gather(G, 2, a)                                                     # (R, M, 1)
        Gaa = torch.gather(Gia, 1, self.slot.view(1, 1, 1).expand(R, 1, 1))            # (R, 1, 1)
        Gd = G - Gia - Gia.transpose(1, 2) + Gaa
        r = (Gaa - Gia)[:, :, 0]
        valid = (self.slots < self.count) & (self.slots != self.slot)                  # (M,)
        keep = valid[None, :, None] & valid[None, None, :]
        Gs = torch.where(keep, Gd, self.eye_m.expand(R, M, M))
        rs = torch.where(valid[None, :], r, torch.zeros_like(r))
        self.G.copy_(Gs)
        self.rhs.copy_(rs)
        _lib.check(L.aimdhf_batch_diis_solve(R, M, _ptr(self.G), _ptr(self.rhs), _ptr(self.coef), _stream()))
        c = torch.where(valid[None, :], self.coef, torch.zeros_like(self.coef))
        anchor = (self.slots[None, :] == self.slot[:, None]).to(c.dtype)                # (1, M)
        c = c + anchor * (1.0 - c.sum(dim=1, keepdim=True))
        bad = ~(torch.isfinite(c).all(dim=1) & (c.abs().amax(dim=1) <= 1e4))
        c = torch.where(bad[:, None], anchor.expand(R, M), c)
        Fx = (c[:, :, None, None, None] * self.Fbuf).sum(dim=1)                         # (R, 2, n, n)
        Fx = Fx + PAD_LEVEL * self.pad_diag[:, None]
        _lib.check(L.aimdhf_batch_purify(R * 2, n, _ptr(Fx.contiguous()), _ptr(self.nocc_t), _ptr(self.Pbuf),
                                         self.purify_steps, 1e-11, _ptr(self.perr), _stream()))
        Dn = Xe @ self.Pbuf.view(R, 2, n, n) @ Xe.transpose(2, 3)
        conv = ((self.energy_el - self.e_prev).abs() < self.conv_tol) & (self.err < self.conv_tol_grad)
        self.frozen |= conv
        self.D.copy_(torch.where(self.frozen[:, None, None, None], D, Dn))
        self.e_prev.copy_(self.energy_el)
        self.slot.copy_((self.slot + 1) % M)
        self.count.copy_(torch.clamp(self.count + 1, max=M))

    def _replay(self) -> None:
        if not self.capture:
            self._iteration()
            return
        if self._graph is None:
            side = torch.cuda.Stream(device=self.device)
            side.wait_stream(torch.cuda.current_stream(self.device))
            saved = [t.clone() for t in self._state()]
            with torch.cuda.stream(side):
                self._iteration()                         # warm-up (libraries allocate workspaces)
            torch.cuda.current_stream(self.device).wait_stream(side)
            for t, v in zip(self._state(), saved):
                t.copy_(v)
            g = torch.cuda.CUDAGraph()
            with torch.cuda.graph(g):
                self._iteration()
            self._graph = g
        self._graph.replay()

    def _state(self):
        return [self.D, self.F, self.Fbuf, self.Ebuf, self.energy_el, self.e_prev, self.err, self.frozen, self.slot,
                self.count, self.J, self.K, self.Pbuf, self.perr, self.G, self.rhs, self.coef]

    def _guess(self) -> torch.Tensor:
        Ds = torch.zeros((self.nrep, 2, self.ld, self.ld), dtype=torch.float64, device=self.device)
        sad: dict[tuple, np.ndarray] = {}
        for r, b in enumerate(self.bases):
            key = tuple(b.symbols)
            if key not in sad:
                sad[key] = sad_density(b)
            P = torch.as_tensor(sad[key], dtype=torch.float64, device=self.device)
            n = b.nao
            Ds[r, 0, :n, :n] = 0.5 * P
            Ds[r, 1, :n, :n] = 0.5 * P
        return Ds

    def scf(self, positions, guess: torch.Tensor | None = None) -> dict:
        """SCF of every replica at ``positions``; the guess is the previous densities
        (or SAD on the first call) unless given."""
        self._set_geometry(positions)
        if self.jk_lanes is None:
            self.autotune()
        D0 = guess if guess is not None else (self.Ds if self.Ds is not None else self._guess())
        self.D.copy_(D0)
        self.frozen.zero_()
        self.e_prev.fill_(float("nan"))
        self.slot.zero_()
        self.count.zero_()
        its = 0
        done = False
        while its < self.max_cycles:
            self._replay()
            its += 1
            if bool(self.frozen.all()):
                done = True
                break
        self.iterations = its
        self.Ds = self.D.clone()
        conv = self.frozen.cpu().numpy()
        return {"energy": self.energy_el.cpu().numpy() + self.e_nuc, "converged": conv, "iterations": its,
                "all_converged": done}

    def gradient(self) -> np.ndarray:
        """(natm_total, 3) analytic gradient of the last SCF of every replica."""
        L = _lib.lib()
        D, F = self.D, self.F
        W = (D @ F @ D).sum(dim=1)
        W = 0.5 * (W + W.transpose(1, 2))
        P = (D[:, 0] + D[:, 1]).contiguous()
        g1 = torch.empty((self.natm, 3), dtype=torch.float64, device=self.device)
        g2 = torch.empty_like(g1)
        _lib.check(L.aimdhf_grad_one(self._h, _ptr(P), _ptr(W.contiguous()), _ptr(g1), _stream()))
        _lib.check(L.aimdhf_grad_two(self._h, _ptr(P), 2, _ptr(D), 1.0, self.gradient_screening, _ptr(g2),
                                     _stream()))
        g = (g1 + g2).cpu().numpy()
        for a, b in zip(self.rep_atom[:-1], self.rep_atom[1:]):
            g[a:b] += nuclear_repulsion_gradient(self.charges[a:b], self.positions[a:b])
        return g

    def compute(self, positions) -> dict:
        """Energies (nrep,), gradient (natm_total, 3), per-replica convergence."""
        out = self.scf(positions)
        out["gradient"] = self.gradient()
        return out

    def densities(self) -> list[np.ndarray]:
        """Per replica: the total density (RHF) or the (2, nao, nao) spin densities (UHF),
        the layout of the single-molecule backend's density."""
        D = self.D.cpu().numpy()
        out = []
        for r, n in enumerate(self.rep_nao):
            Ds = D[r, :, :n, :n]
            out.append(Ds.copy() if self.unrestricted[r] else Ds[0] + Ds[1])
        return out


class BatchLangevinBAOAB:
    """
    Langevin BAOAB (aimd/integrators.py:450-471) for every replica of a batch on
    device tensors: B v += dt/2 a; A x += dt/2 v; O v = c1 v + c2 sqrt(kT/m) xi;
    A x += dt/2 v; forces at x; B v += dt/2 a. ``heat`` (per replica) accumulates
    the kinetic energy the O step changes. The noise comes from a seeded torch
    generator on the device.
    """

    def __init__(self, batch: ReplicaBatch, timestep_fs: float, temperature_k: float,
                 friction_per_fs: float = 0.001, seed: int = 0) -> None:
        self.batch = batch
        dev = batch.device
        self.dt = float(timestep_fs) * FS_TO_AU_TIME
        self.kT = KB_AU * float(temperature_k)
        gamma = float(friction_per_fs) / FS_TO_AU_TIME
        self.c1 = math.exp(-gamma * self.dt)
        self.c2 = math.sqrt(1.0 - self.c1 ** 2)
        self.m = torch.tensor(batch.masses, dtype=torch.float64, device=dev)[:, None]
        self.rep_of_atom = torch.repeat_interleave(torch.arange(batch.nrep, device=dev),
                                                   torch.tensor(batch.natm_rep, device=dev))
        self.gen = torch.Generator(device=dev)
        self.gen.manual_seed(int(seed))
        self.heat = torch.zeros(batch.nrep, dtype=torch.float64, device=dev)
        self.result = None

    def _ekin(self, v: torch.Tensor) -> torch.Tensor:
        e = 0.5 * (self.m * v * v).sum(dim=1)
        return torch.zeros_like(self.heat).index_add_(0, self.rep_of_atom, e)

    def initialize(self, x: torch.Tensor, v: torch.Tensor) -> dict:
        self.x, self.v = x.clone(), v.clone()
        self.result = self.batch.compute(self.x.cpu().numpy())
        self.a = -torch.as_tensor(self.result["gradient"], device=self.x.device) / self.m
        return self.result

    def step(self) -> dict:
        half = 0.5 * self.dt
        self.v += half * self.a
        self.x += half * self.v
        before = self._ekin(self.v)
        xi = torch.randn(self.v.shape, generator=self.gen, dtype=torch.float64, device=self.v.device)
        self.v = self.c1 * self.v + self.c2 * torch.sqrt(self.kT / self.m) * xi
        self.heat += self._ekin(self.v) - before
        self.x += half * self.v
        self.result = self.batch.compute(self.x.cpu().numpy())
        self.a = -torch.as_tensor(self.result["gradient"], device=self.x.device) / self.m
        self.v += half * self.a
        return self.result

    def conserved_energy(self) -> np.ndarray:
        return self.result["energy"] + self._ekin(self.v).cpu().numpy() - self.heat.cpu().numpy()
