"""
Host-side SCF pieces on CPU tensors (no GPU): the DIIS solver against the
linear error model and the blow-up rule of the reference's own tests
(tests/test_qc_scf.py:763-796 of the reference), and option validation.
"""

import numpy as np
import pytest
import torch

from aimd_amd.qc import scf as S


def _sym(n, rng):
    a = rng.standard_normal((n, n))
    return torch.tensor(0.5 * (a + a.T))[None]


def test_diis_is_exact_for_a_linear_error_model():
    rng = np.random.default_rng(4)
    F_star = _sym(5, rng)
    G = _sym(5, rng)
    d = S.DIIS(4)
    d.push(F_star + G, G)                 # error proportional to F - F*
    d.push(F_star - 0.5 * G, -0.5 * G)
    np.testing.assert_allclose(d.extrapolate().numpy(), F_star.numpy(), atol=1e-12)
    d.push(F_star - 0.5 * G, -0.5 * G)    # duplicate -> singular system, handled
    out = d.extrapolate().numpy()
    assert np.all(np.isfinite(out))
    np.testing.assert_allclose(out, F_star.numpy(), atol=1e-10)
    for _ in range(6):
        d.push(F_star, torch.zeros_like(G))
    assert len(d) == 4                    # bounded history
    np.testing.assert_allclose(d.extrapolate().numpy(), F_star.numpy(), atol=1e-12)


def test_diis_matches_least_squares_on_random_vectors():
    # anchored Pulay DIIS = min |e_m + sum_i c_i (e_i - e_m)| (numpy lstsq), c_m = 1 - sum c_i
    rng = np.random.default_rng(11)
    d = S.DIIS(6)
    errs = [_sym(7, rng) for _ in range(5)]
    for e in errs:
        d.push(_sym(7, rng), e)
    c = d.coefficients()
    E = np.stack([e.numpy().ravel() for e in errs])
    A = (E[:-1] - E[-1]).T
    x, *_ = np.linalg.lstsq(A, -E[-1], rcond=None)
    np.testing.assert_allclose(c[:-1], x, rtol=1e-9, atol=1e-12)
    assert abs(c.sum() - 1.0) < 1e-12


def test_diis_host_gram_from_the_raw_slot_matrix_equals_the_device_gram():
    # the SCF loop reads the raw (space, space) slot Gram with its scalars and forms the
    # differences on the host; with the ring wrapped (more pushes than slots) and after a
    # restart from the newest vector both must equal gram() bit for bit
    rng = np.random.default_rng(12)
    d = S.DIIS(4)
    for k in range(7):
        d.push(_sym(6, rng), _sym(6, rng))
        host = d.gram_from(d.raw_gram.numpy().copy())
        dev = d.gram()
        assert (host is None) == (dev is None)
        if dev is not None:
            np.testing.assert_array_equal(host, dev.numpy())
    d.keep_last()
    assert d.gram_from(d.raw_gram.numpy().copy()) is None
    d.push(_sym(6, rng), _sym(6, rng))
    np.testing.assert_array_equal(d.gram_from(d.raw_gram.numpy().copy()), d.gram().numpy())


def test_diis_drops_the_oldest_vectors_when_the_extrapolation_blows_up():
    rng = np.random.default_rng(7)
    u, w = _sym(4, rng), _sym(4, rng)
    F_old, F_new = _sym(4, rng), _sym(4, rng)
    d = S.DIIS(4)
    d.push(F_old, u + 1e-7 * w)
    d.push(F_new, u)
    assert np.abs(d.coefficients()).max() > d.max_coefficient
    np.testing.assert_array_equal(d.extrapolate().numpy(), F_new.numpy())
    assert len(d) == 1


@pytest.mark.parametrize("n,nocc,gap", [(30, 7, 0.3), (60, 25, 0.05), (40, 0, 1.0), (40, 40, 1.0)])
def test_purified_projector_matches_the_eigenvector_projector(n, nocc, gap):
    rng = np.random.default_rng(n + nocc)
    w = np.sort(rng.uniform(-20.0, 5.0, n))
    w[nocc:] += gap + (w[nocc - 1] - w[nocc] if 0 < nocc < n else 0.0)
    Q, _ = np.linalg.qr(rng.standard_normal((n, n)))
    F = torch.tensor((Q * w) @ Q.T)
    P = S.purified_projector(F, nocc)
    assert P is not None
    ref = Q[:, :nocc] @ Q[:, :nocc].T
    np.testing.assert_allclose(P.numpy(), ref, atol=1e-11)
    assert abs(float(torch.trace(P)) - nocc) < 1e-10


def test_purified_projector_reports_a_missing_gap():
    # degenerate HOMO/LUMO: no projector exists, purification must not claim one
    w = np.array([-3.0, -1.0, -1.0, 2.0])
    F = torch.tensor(np.diag(w))
    assert S.purified_projector(F, 2, max_iter=40) is None


def test_newton_schulz_inverse_square_root_matches_eigh():
    rng = np.random.default_rng(5)
    A = rng.standard_normal((40, 40))
    S = A @ A.T / 40 + 0.05 * np.eye(40)
    w, U = np.linalg.eigh(S)
    ref = (U / np.sqrt(w)) @ U.T
    out = S_module().inverse_sqrt_newton_schulz(torch.tensor(S))
    assert out is not None
    X, smin = out
    np.testing.assert_allclose(X.numpy(), ref, atol=1e-11)
    # 1 / (largest absolute row sum of S^-1): a lower bound of the smallest eigenvalue, within
    # the factor sqrt(n) by which a row sum can exceed the spectral norm
    assert w[0] / np.sqrt(40) <= smin <= w[0] * (1 + 1e-12)
    X, n_removed, smin2 = S_module().orthogonalizer(torch.tensor(S))
    np.testing.assert_allclose(X.numpy(), ref, atol=1e-11)
    assert n_removed == 0


def test_orthogonalizer_falls_back_to_eigh_near_linear_dependence():
    # smallest eigenvalue 1e-9: below the threshold, symmetric orthogonalisation is refused
    # ("auto" switches to canonical and removes one combination), as the reference does
    rng = np.random.default_rng(6)
    Q, _ = np.linalg.qr(rng.standard_normal((12, 12)))
    w = np.linspace(1e-9, 2.0, 12)
    S = (Q * w) @ Q.T
    X, n_removed, smin = S_module().orthogonalizer(torch.tensor(S), "auto", 1e-7)
    assert n_removed == 1 and X.shape == (12, 11)
    with pytest.raises(ValueError):
        S_module().orthogonalizer(torch.tensor(S), "symmetric", 1e-7)


def S_module():
    return S


@pytest.mark.parametrize("bad", [dict(max_iter=0), dict(conv_energy=0.0), dict(guess="minao"), dict(diis_space=1),
                                 dict(damping=1.0), dict(level_shift=-0.1), dict(orthogonalization="lowdin"),
                                 dict(blas_threads=0), dict(jk_mode="tensor"), dict(max_store_bytes=-1),
                                 dict(density_solver="lanczos"), dict(incremental_threshold=-1.0),
                                 dict(incremental_threshold=float("nan"))])
def test_option_validation(bad):
    with pytest.raises(ValueError):
        S.SCFOptions(**bad)


def test_incremental_fock_defaults_follow_the_precision():
    class Solver:
        incremental = S.SCFSolver.incremental

        def __init__(self, **kw):
            self.options = S.SCFOptions(**kw)

    assert not Solver().incremental
    assert Solver(precision="mixed").incremental
    assert Solver(incremental_fock=True).incremental
    assert not Solver(precision="mixed", incremental_fock=False).incremental
