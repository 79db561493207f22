"""
The batched-replica kernels on their own: the purification (X^2 on the FP64
matrix cores) against the exact aufbau projector of random gapped Fock
matrices, and the batched J/K digestion at every lanes-per-element choice
against the single-molecule engine's J/K of each replica.
"""

import ctypes
import os

import numpy as np
import pytest

from tests.conftest import ANG, ROOT

pytestmark = pytest.mark.gpu


def _ptr(t):
    return ctypes.c_void_p(t.data_ptr())


def _stream():
    import torch
    return ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)


@pytest.mark.parametrize("ld", [7, 20, 50, 64])
def test_batched_purification_is_the_aufbau_projector(ld):
    import torch
    from aimd_amd import _lib
    rng = np.random.default_rng(ld)
    nmat = 6
    F = np.zeros((nmat, ld, ld))
    Pref = np.zeros_like(F)
    nocc = rng.integers(1, ld, nmat).astype(float)
    for i in range(nmat):
        U, _ = np.linalg.qr(rng.standard_normal((ld, ld)))
        k = int(nocc[i])
        eps = np.concatenate([rng.uniform(-2.0, -0.4, k), rng.uniform(0.3, 2.5, ld - k)])
        F[i] = (U * eps) @ U.T
        Pref[i] = U[:, :k] @ U[:, :k].T
    dev = "cuda"
    Ft = torch.tensor(F, device=dev)
    nt = torch.tensor(nocc, device=dev)
    P = torch.zeros_like(Ft)
    err = torch.zeros(nmat, dtype=torch.float64, device=dev)
    _lib.check(_lib.lib().aimdhf_batch_purify(nmat, ld, _ptr(Ft), _ptr(nt), _ptr(P), 60, 1e-11, _ptr(err), _stream()))
    Pg = P.cpu().numpy()
    dev_max = np.abs(Pg - Pref).max()
    print(f"ld {ld}: max |P - projector| {dev_max:.2e}, idempotency errors {err.cpu().numpy().max():.1e}", flush=True)
    assert dev_max < 1e-10
    np.testing.assert_allclose(np.trace(Pg, axis1=1, axis2=2), nocc, rtol=0, atol=1e-10)


def test_batched_jk_at_every_lane_choice_matches_the_single_molecule_engine():
    import torch
    from aimd_amd import _lib
    from aimd_amd.batch import ReplicaBatch
    from aimd_amd.md import read_xyz
    from aimd_amd.qc.basis import build_basis
    from aimd_amd.qc.engine import Engine
    names = ["water.xyz", "methanol.xyz", "h4_cluster.xyz", "methyl_radical.xyz"]
    mols, xs = [], []
    for f in names:
        s, x = read_xyz(os.path.join(ROOT, "tests", "golden", "replicas", f))
        mols.append((s, 0, 2 if f.startswith("methyl") else 1))
        xs.append(x * ANG)
    batch = ReplicaBatch(mols, basis="6-31g**", jk_lanes=64)
    batch._set_geometry(np.concatenate(xs))
    rng = np.random.default_rng(4)
    R, n = batch.nrep, batch.ld
    D = torch.zeros((R, 2, n, n), dtype=torch.float64, device="cuda")
    for r, nr in enumerate(batch.rep_nao):
        A = rng.standard_normal((2, nr, nr)) * 0.1
        D[r, :, :nr, :nr] = torch.tensor(A @ A.transpose(0, 2, 1), device="cuda")
    out = {}
    for lanes in (64, 32, 16):
        _lib.check(_lib.lib().aimdhf_batch_set_jk_lanes(batch._h, lanes))
        batch._jk(D)
        out[lanes] = (batch.J.clone(), batch.K.clone())
    for lanes in (32, 16):
        assert float((out[lanes][0] - out[64][0]).abs().max()) < 1e-12
        assert float((out[lanes][1] - out[64][1]).abs().max()) < 1e-12
    for r, ((s, _, _), x) in enumerate(zip(mols, xs)):
        b = build_basis(s, x, "6-31g**")
        e = Engine(b)
        e.set_geometry(x)
        nr = b.nao
        J1, K1 = e.jk(D[r, :, :nr, :nr].contiguous())
        for lanes, (J, K) in out.items():
            assert float((J[r, :, :nr, :nr] - J1).abs().max()) < 1e-11, (r, lanes)
            assert float((K[r, :, :nr, :nr] - K1).abs().max()) < 1e-11, (r, lanes)
        e.close()
    times = batch.autotune()
    print("batched J/K ms by lanes per element:", {k: round(v, 4) for k, v in times.items()}, flush=True)
    assert batch.jk_lanes in (16, 32, 64) and batch.jk_lanes == min(times, key=times.get)
    batch.close()
