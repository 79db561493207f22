#!/usr/bin/env python3
"""
Benchmark: ms per Born-Oppenheimer MD step (SCF + forces) of (H2O)16 RHF/6-31G*
(304 Cartesian AOs) on one MI355X.

A step is one velocity-Verlet NVE step at 0.5 fs from 300 K Maxwell-Boltzmann
velocities (seed 1, COM removed): half kick, drift, one ``hf`` backend
compute (device SCF from the previous density + analytic gradient via the
autograd backward), half kick. W untimed warmup steps, then K timed steps
bracketed by torch.cuda.synchronize(). The geometry is the committed fixture
tests/golden/water16.xyz (16 randomly rotated waters in an 8 A cube, O-O >=
2.7 A, seed 1).

By default the FP64 trajectory runs first; the mixed-precision path (FP32 for
quartets below the Schwarz switches, and incremental Fock builds whose
density differences are screened and digested in FP32) then computes the same geometries with
its own density reuse, and its energies and forces are compared with the
FP64 ones at every step. ``value`` is the mixed step time when it stays
within |dE| <= 1e-6 Eh and |dF| <= 1e-5 Eh/bohr at every step, else FP64.

Printed: one JSON line with the step times, the roofline of the kernel with
the largest time per step (HIP events on the engine stream in one extra step
with every launch serialised: a derivative-ERI class priced in FP64 flops from
aimd_amd.qc.flopmodel against the FP64 vector peak, or a store-mode digestion
kernel priced in the bytes it streams against HBM), the store's streamed and
unique integral bytes, and a CPU baseline: one whole BOMD step of the oracle
(SCF from the previous step's density plus the gradient) timed on the host
cores, with the CPU model.

    python bench.py [--gpus 1] [--steps 10] [--warmup 2] [--precision both|fp64|mixed] [--no-cpu-baseline]

Two further workloads print their own JSON line (not the default bench line):

    --workload water32-xlbomd   (H2O)32 RHF/6-31G* (608 AOs) XL-BOMD (K=5) NVE at 0.5 fs: the FP64
                                trajectory, then one mixed-precision trajectory per FP32 Schwarz
                                switch (1e-5, 1e-7, 1e-9), each step's energy and forces compared
                                with an FP64 SCF at the same geometry; drift of every trajectory
    --workload replicas         128 replicas of water / methanol / H4 / methyl radical (UHF) in
                                6-31G**, Langevin BAOAB in one batched plan, FP64 and mixed
"""

from __future__ import annotations

import argparse
import json
import os
import sys
import time

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(HERE, "ab-initio-md_amd"))

METRIC = "ms per BOMD step (SCF+forces), (H2O)16/6-31G*: FP64 vs mixed at ΔE≤1µEh"
PEAK_FP64_TFLOPS = 78.6      # MI355X FP64 vector (and matrix) peak
PEAK_HBM_GBS = 8000.0


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-threads", type=int, default=0)
    ap.add_argument("--jk-mode", choices=["auto", "direct", "store"], default="auto")
    ap.add_argument("--precision", choices=["both", "fp64", "mixed"], default="both",
                    help="both: FP64 trajectory, then the mixed path over the same geometries")
    ap.add_argument("--serial", action="store_true",
                    help="keep every launch on one stream in the timed region too (for kernel traces)")
    ap.add_argument("--copy-probe-mib", type=int, default=0,
                    help="only run one device-to-device copy of this many MiB (HBM counter calibration)")
    ap.add_argument("--workload", choices=["water16", "water32-xlbomd", "replicas"], default="water16")
    ap.add_argument("--md-steps", type=int, default=0,
                    help="trajectory length of the water32-xlbomd (default 200) and replicas (default 100) workloads")
    ap.add_argument("--replicas", type=int, default=128)
    ap.add_argument("--check-every", type=int, default=1,
                    help="water32-xlbomd: compare every n-th mixed step with an FP64 SCF (the last step always)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU only (got --gpus {args.gpus}); the engine has no multi-GPU path "
              "in this build", file=sys.stderr)
        return 2
    if args.workload == "water32-xlbomd":
        return bench_water32_xlbomd(args)
    if args.workload == "replicas":
        return bench_replicas(args)

    import numpy as np
    import torch

    if args.copy_probe_mib > 0:
        n = args.copy_probe_mib * (1 << 20) // 8
        a = torch.ones(n, dtype=torch.float64, device="cuda")
        b = torch.empty_like(a)
        torch.cuda.synchronize()
        b.copy_(a)
        torch.cuda.synchronize()
        print(json.dumps({"copy_probe_bytes": 8 * n}))
        return 0

    from aimd_amd.backends import get_backend
    from aimd_amd.md import VelocityVerletNVE, maxwell_boltzmann, masses_au, read_xyz
    from aimd_amd.qc import flopmodel
    from aimd_amd.units import ANG_TO_BOHR

    sym, xyz = read_xyz(os.path.join(HERE, "tests", "golden", "water16.xyz"))
    x0 = xyz * ANG_TO_BOHR
    m = masses_au(sym)
    v0 = maxwell_boltzmann(m, 300.0, seed=1)

    def backend(precision):
        return get_backend("hf")(sym, basis="6-31g*", conv_tol=1e-10, conv_tol_grad=1e-7,
                                 scf_options={"jk_mode": args.jk_mode, "precision": precision})

    def timed(engine, fn):
        torch.cuda.synchronize()
        engine.profile(True)
        engine.profile_read(reset=True)
        t0 = time.perf_counter()
        out = fn()
        torch.cuda.synchronize()
        wall = time.perf_counter() - t0
        ms, n = engine.profile_read(reset=True)
        engine.profile(False)
        return out, 1e3 * wall / args.steps, ms, n

    # FP64 reference trajectory: velocity-Verlet steps, geometries and results recorded
    precisions = ["fp64", "mixed"] if args.precision == "both" else [args.precision]
    be = backend(precisions[0])
    md = VelocityVerletNVE(be, m, x0, v0, timestep_fs=0.5)
    eng = be.solver.engine
    if args.serial:
        eng.set_streams(False)
    md.initialize()
    warm_x = []
    for _ in range(args.warmup):
        md.step()
        warm_x.append(md.x.copy())

    def run_md():
        rec = []
        for _ in range(args.steps):
            r = md.step()
            rec.append((md.x.copy(), r, md.total_energy()))
        return rec

    rec, ms_per_step, ms_step_kind, n_step_kind = timed(eng, run_md)
    # per-kernel figures (roofline): one more step with every launch serialised on the engine
    # stream, so each kernel's HIP-event time is its own and not shared with overlapping ones
    P_prev = be.last_scf.density.copy()
    eng.set_streams(False)
    steps_timed = args.steps
    args.steps = 1
    _, _, ms_kind, n_kind = timed(eng, lambda: md.step())
    args.steps = steps_timed
    eng.set_streams(not args.serial)
    iters = [r.info["scf_iterations"] for _, r, _ in rec]
    tim = [r.info["timings"] for _, r, _ in rec]
    e_tot = [e for _, _, e in rec]
    mixed = None
    if len(precisions) == 2:
        # mixed precision over the same geometries (same initial point, warm-up and timed steps)
        bm = backend("mixed")
        em = bm.solver.engine
        bm.compute(x0)
        for xw in warm_x:
            bm.compute(xw)
        outs, mixed_ms, _, _ = timed(em, lambda: [bm.compute(xr) for xr, _, _ in rec])
        de = max(abs(o.energy - r.energy) for o, (_, r, _) in zip(outs, rec))
        dfm = max(float(np.abs(o.gradient - r.gradient).max()) for o, (_, r, _) in zip(outs, rec))
        mixed = {"ms_per_step": round(mixed_ms, 3), "max_abs_dE_Eh": de, "max_abs_dF_Eh_bohr": dfm,
                 "within_bounds": bool(de <= 1e-6 and dfm <= 1e-5),
                 "jk_fp32_switch": bm.solver.options.jk_fp32_switch,
                 "grad_fp32_switch": bm.solver.options.grad_fp32_switch,
                 "delta_fp32_switch": bm.solver.options.delta_fp32_switch,
                 "incremental_fock": bm.solver.incremental,
                 "incremental_threshold": bm.solver.incremental_threshold,
                 "scf_iterations_mean": float(np.mean([o.info["scf_iterations"] for o in outs])),
                 "phase_ms": {k: round(1e3 * float(np.mean([o.info["timings"][k] for o in outs])), 3)
                              for k in outs[0].info["timings"]}}

    # Fock build time per build (J and K kernels, reductions, prologue/epilogue)
    quartets = eng.quartets()
    npp = eng.pair_npp()
    info = eng.jk_info()
    tinfo = eng.jk_type_info()
    per_build_45 = 1 if info["stored"] else 3        # prologue/epilogue launches per build
    n_builds = max(int(n_kind[45]) // per_build_45, 1)
    jk_ms = float(ms_kind[eng.FOCK_KINDS].sum()) / n_builds
    store_ms = float(ms_kind[eng.STORE_KINDS].sum())
    grad_ms = float(ms_kind[eng.GRAD_KINDS].sum())
    gw = flopmodel.grad_work(be.basis, eng.grad_quartets(), npp)
    grad_main_ms = float(ms_kind[21:42].sum())
    per_kernel = {eng.KIND_NAMES[k]: round(float(ms_kind[k]) / max(int(n_kind[k]), 1), 4)
                  for k in range(eng.NKIND) if n_kind[k] > 0}
    roofline = dominant_kernel_roofline(eng, ms_kind, n_kind, tinfo, gw)
    unique_b = int(tinfo["unique_bytes"].sum())
    streamed_b = int(tinfo["k_bytes"].sum() + tinfo["j_bytes"].sum()
                     + 8 * (tinfo["k_entries"].sum() + tinfo["j_entries"].sum()))
    use_mixed = mixed is not None and mixed["within_bounds"]
    value = mixed["ms_per_step"] if use_mixed else ms_per_step
    out = {
        "metric": METRIC,
        "value": round(value, 3),
        "unit": "ms",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(value, 3),
        "higher_is_better": False,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64+f32 (mixed)" if use_mixed else ("f64" if precisions[0] == "fp64" else "f64+f32 (mixed)"),
        "data": "synthetic",
        "config": {
            "workload": "(H2O)16 RHF/6-31G* (304 AOs), velocity-Verlet NVE BOMD step at 0.5 fs, 300 K, "
                        "SCF conv 1e-10 Eh / 1e-7, Schwarz 1e-12, gradient screening 1e-14; value: the mixed "
                        "path (FP32 below the Schwarz switches, incremental Fock with FP32 density-difference "
                        "builds) over the FP64 trajectory's geometries when it "
                        "stays within |dE| <= 1e-6 Eh and |dF| <= 1e-5 Eh/bohr of FP64 at every step, else FP64; "
                        "per-kernel figures from the FP64 run",
            "geometry": "tests/golden/water16.xyz",
            "quartets": int(len(quartets)),
            "gradient_quartets": int(eng.class_counts()[1].sum()),
        },
        "fp64_ms_per_step": round(ms_per_step, 3) if precisions[0] == "fp64" else None,
        "mixed": mixed,
        "mixed_speedup": round(ms_per_step / mixed["ms_per_step"], 3) if mixed else None,
        "scf_iterations_mean": float(np.mean(iters)),
        "phase_ms": {k: round(1e3 * float(np.mean([t[k] for t in tim])), 3) for k in tim[0]} if tim else {},
        "jk_mode": "store" if info["stored"] else "direct",
        "jk_builds_per_step": round(n_builds, 2),
        "kernel_figures": "jk_*, store_*, gradient_*, roofline and avg_ms_per_launch come from one extra step "
                          "with every launch serialised on the engine stream after the timed region",
        "jk_build_ms": round(float(jk_ms), 3),
        "jk_traffic": {"unique_integral_bytes": unique_b, "streamed_bytes_per_build": streamed_b,
                       "streamed_over_unique": round(streamed_b / unique_b, 3) if unique_b else None,
                       "stream_gbs": round(streamed_b / (jk_ms * 1e-3) / 1e9, 1) if jk_ms > 0 and unique_b else None,
                       "note": "streamed = gather-ordered integral copies plus 8-byte list entries that one "
                               "J/K build reads; unique = the screened integrals stored once"},
        "store_build_ms_per_step": round(store_ms, 3),
        "store_bytes": info["store_bytes"],
        "gradient_kernels_ms_per_step": round(float(grad_ms), 3),
        "gradient_tflops": round(gw["total"] / (grad_main_ms * 1e-3) / 1e12, 3) if grad_main_ms > 0 else None,
        "digested_quartets_per_s": float(len(quartets) / (jk_ms * 1e-3)) if jk_ms > 0 else None,
        "nve_total_energy_span_Eh": float(np.ptp(e_tot)) if e_tot else None,
        "roofline": roofline,
        "avg_ms_per_launch": per_kernel,
    }
    if not args.no_cpu_baseline:
        out["cpu_baseline"] = cpu_baseline(sym, md.x, P_prev, args.cpu_threads)
    print(json.dumps(out))
    return 0


def _progress(msg):
    print(msg, file=sys.stderr, flush=True)


def _drift(e, dt_fs):
    """(straight-line drift in Eh/ps, max |E - E0|) of a total-energy series."""
    import numpy as np
    e = np.asarray(e, dtype=float)
    t_ps = np.arange(len(e)) * dt_fs * 1e-3
    return float(np.polyfit(t_ps, e, 1)[0]), float(np.abs(e - e[0]).max())


def bench_water32_xlbomd(args):
    """(H2O)32 RHF/6-31G* XL-BOMD: FP64 trajectory and the FP32-switch sweep (module docstring)."""
    import numpy as np
    import torch
    from aimd_amd import md
    from aimd_amd.backends import get_backend
    from aimd_amd.units import ANG_TO_BOHR

    steps = args.md_steps or 200
    sym, xyz = md.read_xyz(os.path.join(HERE, "tests", "golden", "water32.xyz"))
    x0 = xyz * ANG_TO_BOHR
    v0 = md.maxwell_boltzmann(md.masses_au(sym), 300.0, seed=1)

    def backend(precision, switch=1e-5):
        opts = {"precision": precision, "jk_fp32_switch": switch, "grad_fp32_switch": switch}
        return get_backend("hf")(sym, basis="6-31g*", conv_tol=1e-10, conv_tol_grad=1e-7, scf_options=opts)

    def trajectory(be, label, record=False):
        """Timed XL-BOMD run; with record, every step's geometry, density and result are
        kept on the host for the FP64 checks made after this backend is closed."""
        sysm = md.System.create(sym, x0, v0)
        xl = md.XLBOMD(be, 0.5, k=5)
        xl.initialize(sysm)
        e_tot, its, rec = [xl.conserved_energy(sysm)], [], []
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for k in range(steps):
            r = xl.step(sysm)
            e_tot.append(xl.conserved_energy(sysm))
            its.append(r.info["scf_iterations"])
            if record:
                rec.append((sysm.positions.copy(), np.asarray(r.density), r.energy, np.asarray(r.gradient)))
            if (k + 1) % 20 == 0:
                _progress(f"{label}: step {k + 1}/{steps}, E_tot {e_tot[-1]:.8f}")
        torch.cuda.synchronize()
        ms = 1e3 * (time.perf_counter() - t0) / steps
        be.close()
        drift, span = _drift(e_tot, 0.5)
        return {"ms_per_step": round(ms, 2), "drift_Eh_per_ps": drift, "max_abs_dEtot_Eh": span,
                "scf_iterations_mean": float(np.mean(its))}, rec

    fp64, _ = trajectory(backend("fp64"), "fp64")
    sweep, recs = {}, {}
    for sw in (1e-5, 1e-7, 1e-9):
        sweep[f"{sw:g}"], recs[f"{sw:g}"] = trajectory(backend("mixed", sw), f"mixed {sw:g}", record=True)
    # per-step errors of every mixed trajectory: an FP64 SCF at the same geometry, started
    # from the mixed density (the timed mixed runs above shared the GPU with nothing else)
    check = backend("fp64")
    every = max(int(args.check_every), 1)
    for key, rec in recs.items():
        de, dfm = [], []
        for k, (x, dens, e, g) in enumerate(rec):
            if k % every and k != len(rec) - 1:
                continue
            check.set_density_guess(dens)
            ref = check.compute(x)
            de.append(abs(e - ref.energy))
            dfm.append(float(np.abs(g - ref.gradient).max()))
        _progress(f"checked mixed {key}: {len(de)} steps, max |dE| {max(de):.2e}")
        sweep[key].update({"max_abs_dE_Eh": max(de), "max_abs_dF_Eh_bohr": max(dfm), "steps_checked": len(de),
                           "within_bounds": bool(max(de) <= 1e-6 and max(dfm) <= 1e-5)})
    check.close()
    best = min((v["ms_per_step"], k) for k, v in sweep.items() if v["within_bounds"]) if any(
        v["within_bounds"] for v in sweep.values()) else None
    out = {"metric": "ms per XL-BOMD step (SCF+forces), (H2O)32/6-31G*: FP64 vs mixed FP32-switch sweep",
           "value": best[0] if best else fp64["ms_per_step"], "unit": "ms", "n_gpus": 1, "steps": steps,
           "higher_is_better": False, "dtype": "f64+f32 (mixed)" if best else "f64", "data": "synthetic",
           "config": {"workload": "(H2O)32 RHF/6-31G* (608 AOs), XL-BOMD K=5 NVE at 0.5 fs from 300 K "
                                  "Maxwell-Boltzmann velocities (seed 1), SCF 1e-10 Eh / 1e-7; per mixed step an "
                                  "FP64 SCF at the same geometry gives the energy and force error",
                      "geometry": "tests/golden/water32.xyz", "best_switch": best[1] if best else None},
           "fp64": fp64, "mixed_sweep": sweep}
    print(json.dumps(out))
    return 0


def bench_replicas(args):
    """Batched replicas under Langevin BAOAB (module docstring)."""
    import numpy as np
    import torch
    from aimd_amd.batch import BatchLangevinBAOAB, ReplicaBatch
    from aimd_amd.md import maxwell_boltzmann, read_xyz
    from aimd_amd.units import ANG_TO_BOHR

    steps = args.md_steps or 100
    kinds = [("water.xyz", 0, 1), ("methanol.xyz", 0, 1), ("h4_cluster.xyz", 0, 1), ("methyl_radical.xyz", 0, 2)]
    rng = np.random.default_rng(1)
    pick = rng.integers(0, len(kinds), args.replicas)
    mols, xs = [], []
    for k in pick:
        f, c, mult = kinds[k]
        s, x = read_xyz(os.path.join(HERE, "tests", "golden", "replicas", f))
        mols.append((s, c, mult))
        xs.append((x + rng.normal(0.0, 0.05, x.shape)) * ANG_TO_BOHR)
    x0 = np.concatenate(xs)
    results = {}
    geoms = None
    for precision in ("fp64", "mixed"):
        batch = ReplicaBatch(mols, basis="6-31g**", precision=precision)
        dev = batch.device
        v0 = np.concatenate([maxwell_boltzmann(batch.masses[a:b], 300.0, seed=int(r))
                             for r, (a, b) in enumerate(zip(batch.rep_atom[:-1], batch.rep_atom[1:]))])
        if precision == "fp64":
            integ = BatchLangevinBAOAB(batch, 0.5, 300.0, friction_per_fs=0.001, seed=1)
            integ.initialize(torch.tensor(x0, device=dev), torch.tensor(v0, device=dev))
            for _ in range(max(args.warmup, 1)):
                integ.step()
            torch.cuda.synchronize()
            geoms, energies, grads, iters = [], [], [], []
            t0 = time.perf_counter()
            for k in range(steps):
                r = integ.step()
                geoms.append(integ.x.cpu().numpy())
                energies.append(r["energy"])
                grads.append(r["gradient"])
                iters.append(r["iterations"])
                if (k + 1) % 20 == 0:
                    _progress(f"replicas fp64: step {k + 1}/{steps}")
            torch.cuda.synchronize()
            wall = time.perf_counter() - t0
            results["fp64"] = {"ms_per_step": round(1e3 * wall / steps, 3),
                               "replica_steps_per_s": round(batch.nrep * steps / wall, 1),
                               "scf_iterations_mean": float(np.mean(iters)),
                               "jk_lanes": batch.jk_lanes,
                               "jk_ms_by_lanes": {str(k): round(v, 4) for k, v in batch.jk_times.items()}}
        else:
            batch.compute(x0)
            torch.cuda.synchronize()
            de = np.zeros(batch.nrep)
            dfm = np.zeros(batch.nrep)
            t0 = time.perf_counter()
            for k, g in enumerate(geoms):
                r = batch.compute(g)
                de = np.maximum(de, np.abs(r["energy"] - energies[k]))
                d = np.abs(r["gradient"] - grads[k]).max(axis=1)
                dfm = np.maximum(dfm, [d[a:b].max() for a, b in zip(batch.rep_atom[:-1], batch.rep_atom[1:])])
            torch.cuda.synchronize()
            wall = time.perf_counter() - t0
            results["mixed"] = {"ms_per_step": round(1e3 * wall / steps, 3),
                                "replica_steps_per_s": round(batch.nrep * steps / wall, 1),
                                "max_abs_dE_Eh": float(de.max()), "max_abs_dF_Eh_bohr": float(dfm.max()),
                                "per_replica_max_abs_dE_Eh": [float(v) for v in de],
                                "within_bounds": bool(de.max() <= 1e-6 and dfm.max() <= 1e-5)}
        batch.close()
    use = results["mixed"] if results["mixed"]["within_bounds"] else results["fp64"]
    counts = {f: int((pick == i).sum()) for i, (f, _, _) in enumerate(kinds)}
    out = {"metric": f"replica-steps per second, {args.replicas} replicas, 6-31G**, Langevin BAOAB",
           "value": use["replica_steps_per_s"], "unit": "replica-steps/s", "n_gpus": 1, "steps": steps,
           "higher_is_better": True, "dtype": "f64+f32 (mixed)" if use is results["mixed"] else "f64",
           "data": "synthetic",
           "config": {"workload": f"{args.replicas} replicas drawn by seed 1 from water / methanol / H4 / methyl "
                                  "radical (UHF doublet, others RHF), 6-31G**, iid 0.05 A displacements, 300 K "
                                  "Maxwell-Boltzmann velocities, BAOAB gamma 1/ps, 0.5 fs; one batched plan; "
                                  "mixed computed on the FP64 trajectory's geometries",
                      "replicas_per_kind": counts},
           "fp64": results["fp64"], "mixed": results["mixed"]}
    print(json.dumps(out))
    return 0


def dominant_kernel_roofline(eng, ms_kind, n_kind, tinfo, gw):
    """Roofline of the kernel with the largest time per step in the serialised profiling
    step, among the kinds that time one kernel per launch: the derivative-ERI classes
    (FP64 VALU, flops from flopmodel.grad_work), and the store-mode exchange and Coulomb
    digestion (HBM, the gather-ordered integral copies and list entries one launch
    streams)."""
    from aimd_amd.qc import flopmodel
    kstore = {55 + t: t for t in range(9)}
    jstore = {49 + c: c for c in range(6)}
    cands = list(range(21, 42)) + list(kstore) + list(jstore)
    k = max(cands, key=lambda i: float(ms_kind[i]))
    avg_ms = float(ms_kind[k]) / max(int(n_kind[k]), 1)
    name = eng.KIND_NAMES[k]
    out = {"kind": name, "ms_per_step": round(float(ms_kind[k]), 4), "launches_per_step": int(n_kind[k]),
           "event_avg_launch_ms": round(avg_ms, 4)}
    if k >= 49:
        # store digestion: 20 back-to-back launches with the last build's arguments between
        # one event pair (the per-launch event pairs of the step add dispatch gaps)
        avg_ms = eng.profile_replay(k, 20)
        out["replay_launches"] = 20
    out["avg_launch_ms"] = round(avg_ms, 4)
    if k < 42:
        c = k - 21
        achieved = gw["per_class"][c] / (avg_ms * 1e-3) / 1e12
        out.update({"bound": "valu", "unit": "TFLOP/s", "peak": PEAK_FP64_TFLOPS,
                    "kernel": f"k_grad_class<{c}, double> (derivative ERIs, quartet class "
                              f"({'/'.join(str(v) for v in flopmodel.class_ls(c))}))",
                    "flops_per_launch": float(gw["per_class"][c]),
                    "note": "FP64 vector-ALU kernel (no MFMA instructions) against MI355X's FP64 vector peak"})
    else:
        if k in kstore:
            t = kstore[k]
            nbytes = int(tinfo["k_bytes"][t] + 8 * tinfo["k_entries"][t])
            kern = f"k_kstore<{t // 3}, {t % 3}, 1> (exchange from the integral store, output blocks " \
                   f"{'spd'[t // 3]}{'spd'[t % 3]})"
        else:
            c = jstore[k]
            nbytes = int(tinfo["j_bytes"][c] + 8 * tinfo["j_entries"][c])
            kern = f"k_jstore<{c}, 1> (Coulomb from the integral store, pair class {eng.PAIR_CLASS_NAMES[c]})"
        achieved = nbytes / (avg_ms * 1e-3) / 1e9
        out.update({"bound": "hbm", "unit": "GB/s", "peak": PEAK_HBM_GBS, "kernel": kern,
                    "bytes_per_launch": nbytes,
                    "note": "bytes = gather-ordered integral copies (8 B FP64, 4 B FP32) and 8-byte list "
                            "entries one launch streams; density reads hit L2 and are not counted"})
    out["achieved"] = round(achieved, 4)
    out["frac"] = round(achieved / out["peak"], 5)
    out["traffic"] = measured_traffic(out["kind"])
    return out


def measured_traffic(kind):
    """HBM bytes per launch of a kernel from its committed rocprofv3 --pmc summary
    (FETCH_SIZE and WRITE_SIZE scaled by a calibration copy), or None if there is none."""
    path = os.path.join(HERE, "profiles", f"{kind}_water16_631gs_pmc.txt")
    if not os.path.exists(path):
        return None
    with open(path, encoding="utf-8") as f:
        for line in f:
            if line.strip().startswith("traffic_bytes_per_launch:"):
                return int(line.split(":")[1])
    return None


def _cpu_model():
    try:
        with open("/proc/cpuinfo", encoding="utf-8") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_baseline(sym, x, P_prev, threads):
    """One BOMD step of the oracle (C++/OpenMP restatement of the reference algorithm) on
    the host cores, timed whole: basis and pairs, the integral-direct SCF from the previous
    step's density to 1e-10 Eh / 1e-7, and the analytic gradient."""
    sys.path.insert(0, HERE)
    from oracle import oracle as O
    n = threads or min(len(os.sched_getaffinity(0)), int(os.environ.get("OMP_NUM_THREADS", "64") or 64))
    used = O.set_threads(n)
    t0 = time.perf_counter()
    b = O.make_basis(sym, x, "6-31g*")
    pr = O.Pairs(b)
    st = O.scf(b, guess=P_prev, conv_energy=1e-10, conv_commutator=1e-7, pairs=pr)
    t1 = time.perf_counter()
    O.scf_gradient(b, st, pr)
    t2 = time.perf_counter()
    return {"value": round(1e3 * (t2 - t0), 1), "unit": "ms", "cores": int(used), "kind": "port",
            "cpu": _cpu_model(),
            "sample": f"one oracle BOMD step of (H2O)16/6-31G* at the last trajectory geometry, timed whole: SCF "
                      f"from the previous step's density ({st.iterations} iterations, integral-direct J/K) "
                      f"{t1 - t0:.2f} s, gradient {t2 - t1:.2f} s"}


if __name__ == "__main__":
    sys.exit(main())
