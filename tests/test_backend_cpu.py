"""Host-side behaviour of the hf backend that needs no GPU: registry, argument validation, MD helpers."""

import numpy as np
import pytest


def test_registry_and_unknown_backend():
    from aimd_amd.backends import get_backend, list_backends
    assert "hf" in list_backends()
    assert get_backend(" HF ").name == "hf"
    with pytest.raises(ValueError, match="Unknown backend"):
        get_backend("nope")


@pytest.mark.parametrize("kw,msg", [
    ({"method": "mp2"}, "Hartree-Fock only"),
    ({"reference": "rohf"}, "unknown reference"),
    ({"reference": "rhf", "method": "uhf"}, "conflicts"),
    ({"guess": "huckel"}, "unknown guess"),
    ({"gradient_screening": -1.0}, "gradient_screening"),
    ({"scf_options": {"conv_energy": 1e-8}}, "conv_tol"),
    ({"scf_options": {"warn_unconverged": True}}, "managed by the backend"),
    ({"scf_options": {"diis_space": 1}}, "diis_space"),
    ({"scf_options": {"bogus": 1}}, "invalid scf_options"),
])
def test_constructor_validation_matches_reference(kw, msg):
    """aimd/backends/hf.py:199-226: bad options are ValueError, raised before any device work."""
    from aimd_amd.backends import get_backend
    with pytest.raises(ValueError, match=msg):
        get_backend("hf")(["H", "H"], **kw)


def test_electron_counts_and_options():
    from aimd_amd.qc.scf import SCFOptions, electron_counts
    assert electron_counts(10, 0, 1, "rhf") == (5, 5)
    assert electron_counts(9, 0, 2, "uhf") == (5, 4)
    with pytest.raises(ValueError):
        electron_counts(9, 0, 1, "uhf")
    with pytest.raises(ValueError, match="closed-shell"):
        electron_counts(9, 0, 2, "rhf")
    with pytest.raises(ValueError):
        SCFOptions(max_iter=0)


def test_coincident_nuclei_refused():
    from aimd_amd.qc.scf import check_nuclear_separation
    with pytest.raises(ValueError, match="coincident"):
        check_nuclear_separation(np.array([[0, 0, 0], [0, 0, 1e-6]]), np.array([1, 1]))


def test_maxwell_boltzmann_velocities():
    from aimd_amd.md import masses_au, maxwell_boltzmann, temperature
    m = masses_au(["O", "H", "H"] * 4)
    v = maxwell_boltzmann(m, 300.0, seed=1)
    assert np.abs((m[:, None] * v).sum(axis=0)).max() < 1e-12
    assert temperature(m, v) == pytest.approx(300.0, rel=1e-12)
    np.testing.assert_array_equal(v, maxwell_boltzmann(m, 300.0, seed=1))


def test_flop_model_counts_water(oracle_lib):
    from aimd_amd.qc import flopmodel
    from aimd_amd.qc.basis import build_basis
    from tests.conftest import WATER, bohr
    s, x = bohr(WATER)
    b = build_basis(s, x, "6-31g*")
    ob = oracle_lib.make_basis(s, x, "6-31g*")
    pr = oracle_lib.Pairs(ob)
    q = pr.quartets(1e-12)
    npp = np.diff(pr.arrays()[0]).astype(np.int32)
    w = flopmodel.jk_work(b, q, npp)
    assert w["quartets"].sum() == len(q)
    assert w["total"] > 0
    assert w["prim_quartets"].sum() == pytest.approx(float((npp[q[:, 0]] * npp[q[:, 1]]).sum()))


def test_gradient_flop_model_hand_counts():
    # H2/STO-3G: pairs 0 = (0,0), 1 = (1,0), 2 = (1,1). An ss|ss pass with the derivative:
    # per primitive quartet (20+3*1) + 12 + 3*(NH(1)-1) + NH(1)*NH(0)*(1+2*1) = 56,
    # per bra primitive pair 6 * 2 * n_cd * box = 6 * 2 * 1 * 8 = 96.
    from aimd_amd.qc import flopmodel
    from aimd_amd.qc.basis import build_basis
    b = build_basis(["H", "H"], np.array([[0.0, 0.0, 0.0], [0.0, 0.0, 1.4]]), "sto-3g")
    npp = np.array([9, 4, 9], np.int32)
    one = 9 * 4 * 56 + 4 * 96            # (1|0): the two-centre pair needs its pass, the ket is one-centre
    w = flopmodel.grad_work(b, np.array([[1, 0]]), npp)
    assert w["total"] == one
    w = flopmodel.grad_work(b, np.array([[2, 0]]), npp)   # both pairs one-centre: one pass
    assert w["total"] == 9 * 9 * 56 + 9 * 96
    w = flopmodel.grad_work(b, np.array([[1, 1]]), npp)   # both two-centre: bra and ket passes
    assert w["total"] == 2 * (4 * 4 * 56 + 4 * 96)
