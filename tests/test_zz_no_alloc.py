"""
The allocation-free one-electron gradient entry point (aimdhf_grad_one_noalloc)
against aimdhf_grad_one through the C ABI: bitwise equal on two calls in a row,
the second on a smaller plan (the fixed scratch keeps the first call's sums,
which must not leak into the second), and a plan beyond the scratch's size is
refused with an error before anything is launched.
"""

import numpy as np
import pytest

from tests.conftest import WATER, bohr, cluster

pytestmark = pytest.mark.gpu


def _plan(s, x, basis):
    import torch
    from aimd_amd.qc.basis import build_basis
    from aimd_amd.qc.engine import Engine
    b = build_basis(s, x, basis)
    e = Engine(b)
    e.set_geometry(x)
    torch.cuda.synchronize()
    return b, e


def _densities(n, seed):
    rng = np.random.default_rng(seed)
    D = rng.standard_normal((n, n)) * 0.1
    W = rng.standard_normal((n, n)) * 0.1
    return D + D.T, W + W.T


def test_noalloc_matches_grad_one_on_two_calls_the_second_smaller():
    import torch
    for name, (s, x), basis, seed in (("water8 6-31g*", cluster("water8.xyz"), "6-31g*", 1),
                                      ("water 6-31g*", bohr(WATER, 0.05, 2), "6-31g*", 2)):
        print(f"case {name}", flush=True)
        b, e = _plan(s, x, basis)
        D, W = _densities(b.nao, seed)
        Dt, Wt = torch.tensor(D, device="cuda"), torch.tensor(W, device="cuda")
        ref = e.grad_one(Dt, Wt)
        out = torch.full((b.natm, 3), np.nan, dtype=torch.float64, device="cuda")
        e.grad_one_noalloc(Dt, Wt, out)
        torch.cuda.synchronize()
        assert torch.equal(out, ref), (out - ref).abs().max().item()
        e.close()


def test_noalloc_refuses_plans_beyond_its_scratch():
    from aimd_amd import _lib
    from aimd_amd.qc.basis import build_basis
    from aimd_amd.qc.engine import Engine
    s, x = cluster("water32.xyz")
    s, x = s * 2, np.concatenate([x, x + 40.0])        # 64 waters, 1216 AOs > 1024
    b = build_basis(s, x, "6-31g*")
    e = Engine(b)
    import torch
    D = torch.zeros((b.nao, b.nao), dtype=torch.float64, device="cuda")
    out = torch.zeros((b.natm, 3), dtype=torch.float64, device="cuda")
    with pytest.raises(ValueError, match="scratch"):
        e.grad_one_noalloc(D, D, out)
    e.close()
