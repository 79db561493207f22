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

Printed: one JSON line with the step time, the dominant kernel's roofline
(algorithmic flops per launch from aimd_amd.qc.flopmodel over the kernel's
average HIP-event duration on the engine stream) and a CPU baseline (the
oracle's integral-direct J/K build and ERI gradient for the same geometry on
the host cores, scaled to a step with the measured mean SCF iteration count).

    python bench.py [--gpus 1] [--steps 10] [--warmup 2] [--no-cpu-baseline]
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
    ap.add_argument("--precision", choices=["fp64", "mixed"], default="fp64")
    ap.add_argument("--copy-probe-mib", type=int, default=0,
                    help="only run one device-to-device copy of this many MiB (HBM counter calibration)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU only (got --gpus {args.gpus}); the engine has no multi-GPU path "
              "in this build", file=sys.stderr)
        return 2

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
    be = get_backend("hf")(sym, basis="6-31g*", conv_tol=1e-10, conv_tol_grad=1e-7,
                           scf_options={"jk_mode": args.jk_mode, "precision": args.precision})
    md = VelocityVerletNVE(be, m, x0, v0, timestep_fs=0.5)
    eng = be.solver.engine
    md.initialize()
    for _ in range(args.warmup):
        md.step()
    torch.cuda.synchronize()
    eng.profile(True)
    eng.profile_read(reset=True)
    iters, e_tot, tim = [], [], []
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        r = md.step()
        iters.append(r.info["scf_iterations"])
        tim.append(r.info["timings"])
        e_tot.append(md.total_energy())
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    ms_kind, n_kind = eng.profile_read(reset=True)
    eng.profile(False)
    ms_per_step = 1e3 * wall / args.steps

    # Fock build (J and K kernels + their prologue/epilogue) priced with the algorithmic
    # work of the unique screened quartets of the last timed geometry
    quartets = eng.quartets()
    npp = eng.pair_npp()
    work = flopmodel.jk_work(be.basis, quartets, npp, nset=1)
    info = eng.jk_info()
    per_build_45 = 1 if info["stored"] else 3        # prologue/epilogue launches per build
    n_builds = max(int(n_kind[45]) // per_build_45, 1)
    fock_ms = float(ms_kind[eng.FOCK_KINDS].sum()) / n_builds
    store_ms = float(ms_kind[eng.STORE_KINDS].sum()) / max(args.steps, 1)
    flops = work["total"]
    jk_ms = fock_ms
    grad_ms = float(ms_kind[eng.GRAD_KINDS].sum()) / max(args.steps, 1)
    # derivative ERIs of the last timed step priced with the gradient work model; the
    # dominant single kernel of the step is the heaviest derivative-ERI class kernel
    gw = flopmodel.grad_work(be.basis, eng.grad_quartets(), npp)
    cls_ms = np.array([ms_kind[21 + k] / max(int(n_kind[21 + k]), 1) for k in range(21)])
    kd = int(np.argmax(cls_ms))
    achieved = gw["per_class"][kd] / (cls_ms[kd] * 1e-3) / 1e12
    grad_main_ms = float(ms_kind[21:42].sum()) / max(args.steps, 1)
    per_kernel = {eng.KIND_NAMES[k]: round(float(ms_kind[k]) / max(int(n_kind[k]), 1), 4)
                  for k in range(eng.NKIND) if n_kind[k] > 0}
    out = {
        "metric": METRIC,
        "value": round(ms_per_step, 3),
        "unit": "ms",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_per_step, 3),
        "higher_is_better": False,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {
            "workload": "(H2O)16 RHF/6-31G* (304 AOs), velocity-Verlet NVE BOMD step at 0.5 fs, 300 K, "
                        "SCF conv 1e-10 Eh / 1e-7, Schwarz 1e-12, gradient screening 1e-14, precision fp64",
            "geometry": "tests/golden/water16.xyz",
            "quartets": int(len(quartets)),
            "gradient_quartets": int(eng.class_counts()[1].sum()),
        },
        "scf_iterations_mean": float(np.mean(iters)),
        "phase_ms": {k: round(1e3 * float(np.mean([t[k] for t in tim])), 3) for k in tim[0]} if tim else {},
        "jk_mode": "store" if info["stored"] else "direct",
        "jk_builds_per_step": round(n_builds / max(args.steps, 1), 2),
        "jk_build_ms": round(float(jk_ms), 3),
        "jk_build_tflops": round(flops / (jk_ms * 1e-3) / 1e12, 3) if jk_ms > 0 else None,
        "jk_stream_gbs": round(info["gather_bytes"] / (jk_ms * 1e-3) / 1e9, 1) if jk_ms > 0 else None,
        "store_build_ms_per_step": round(store_ms, 3),
        "store_bytes": info["store_bytes"],
        "gradient_kernels_ms_per_step": round(float(grad_ms), 3),
        "gradient_tflops": round(gw["total"] / (grad_main_ms * 1e-3) / 1e12, 3) if grad_main_ms > 0 else None,
        "eri_quartets_per_s": float(len(quartets) / (jk_ms * 1e-3)) if jk_ms > 0 else None,
        "nve_total_energy_span_Eh": float(np.ptp(e_tot)) if e_tot else None,
        "roofline": {
            "bound": "mfma",
            "kernel": f"k_grad_class<{kd}> (derivative ERIs, quartet class {kd}: "
                      f"({'/'.join(str(v) for v in flopmodel.class_ls(kd))}), fp64 VALU)",
            "achieved": round(achieved, 4),
            "peak": PEAK_FP64_TFLOPS,
            "unit": "TFLOP/s",
            "frac": round(achieved / PEAK_FP64_TFLOPS, 5),
            "traffic": measured_traffic(kd),
            "flops_per_launch": float(gw["per_class"][kd]),
            "avg_launch_ms": round(float(cls_ms[kd]), 4),
            "note": "FP64 vector kernel priced against MI355X's FP64 peak (78.6 TF/s, vector and matrix)",
        },
        "avg_ms_per_launch": per_kernel,
    }
    if not args.no_cpu_baseline:
        out["cpu_baseline"] = cpu_baseline(sym, md.x, np.mean(iters), args.cpu_threads, be.last_scf.density)
    print(json.dumps(out))
    return 0


def measured_traffic(kernel_class):
    """HBM bytes per launch of k_grad_class<k> from its committed rocprofv3 --pmc summary
    (FETCH_SIZE and WRITE_SIZE scaled by a calibration copy), or None if there is none."""
    path = os.path.join(HERE, "profiles", f"k_grad_class_{kernel_class}_water16_631gs_pmc.txt")
    if not os.path.exists(path):
        return None
    with open(path, encoding="utf-8") as f:
        for line in f:
            if line.strip().startswith("traffic_bytes_per_launch:"):
                return int(line.split(":")[1])
    return None


def cpu_baseline(sym, x, mean_iters, threads, P):
    """The oracle (C++/OpenMP restatement of the reference algorithm) on the host cores:
    one integral-direct J/K build and one ERI gradient of the same geometry, scaled to a
    step as mean_iters * t_jk + t_grad (the one-electron and dense-algebra parts, < 1% of
    the oracle's step, are not included)."""
    import numpy as np
    sys.path.insert(0, HERE)
    from oracle import oracle as O
    n = threads or min(len(os.sched_getaffinity(0)), int(os.environ.get("OMP_NUM_THREADS", "64") or 64))
    used = O.set_threads(n)
    b = O.make_basis(sym, x, "6-31g*")
    t0 = time.perf_counter()
    pr = O.Pairs(b)
    t_pairs = time.perf_counter() - t0
    t0 = time.perf_counter()
    pr.jk_direct(0.5 * P)
    t_jk = time.perf_counter() - t0
    t0 = time.perf_counter()
    pr.eri_gradient(P, [P], 0.5)
    t_grad = time.perf_counter() - t0
    ms = 1e3 * (mean_iters * t_jk + t_grad + t_pairs)
    return {"value": round(ms, 1), "unit": "ms", "cores": int(used), "kind": "port",
            "sample": f"oracle on (H2O)16/6-31G*: pair build + Schwarz {t_pairs:.2f} s, one integral-direct J/K "
                      f"build {t_jk:.2f} s, one ERI gradient {t_grad:.2f} s; step = {mean_iters:.2f} x J/K + "
                      f"gradient + pairs"}


if __name__ == "__main__":
    sys.exit(main())
