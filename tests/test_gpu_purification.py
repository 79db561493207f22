"""
The purification's start and error kernels through the C ABI against numpy: the
Gershgorin start X0 = (e_max - F) / (e_max - e_min) and the idempotency error
|tr X - sum_ij X_ij^2|, at sizes below and above one block of rows, bitwise
repeatable.
"""

import ctypes

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _ptr(t):
    return ctypes.c_void_p(t.data_ptr())


@pytest.mark.parametrize("n", [1, 17, 304, 1000])
def test_tc2_start_and_error_against_numpy(n):
    import torch
    from aimd_amd import _lib
    L = _lib.lib()
    rng = np.random.default_rng(n)
    A = rng.standard_normal((n, n))
    F = A + A.T
    off = np.abs(F).sum(axis=1) - np.abs(np.diag(F))
    emin, emax = (np.diag(F) - off).min(), (np.diag(F) + off).max()
    X0_ref = (emax * np.eye(n) - F) / (emax - emin)
    st = ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)
    Ft = torch.tensor(F, device="cuda")
    X0 = torch.empty_like(Ft)
    _lib.check(L.aimdhf_tc2_init(n, _ptr(Ft), _ptr(X0), st))
    np.testing.assert_allclose(X0.cpu().numpy(), X0_ref, rtol=1e-14, atol=1e-15)
    B = rng.standard_normal((n, n)) * 0.1
    X = 0.5 * (B + B.T)
    err_ref = abs(np.trace(X) - (X * X).sum())
    Xt = torch.tensor(X, device="cuda")
    out = torch.zeros(2, dtype=torch.float64, device="cuda")
    _lib.check(L.aimdhf_tc2_error(n, _ptr(Xt), _ptr(out), st))
    _lib.check(L.aimdhf_tc2_error(n, _ptr(Xt), _ptr(out[1:]), st))
    e = out.cpu().numpy()
    print(f"n {n}: error {e[0]:.6e} vs numpy {err_ref:.6e}", flush=True)
    assert e[0] == e[1]
    assert abs(e[0] - err_ref) <= 1e-13 * max(1.0, (X * X).sum())
