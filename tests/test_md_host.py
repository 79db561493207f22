"""
The MD drivers of aimd_amd.md on the host, with a toy backend whose "SCF" starts
from the density guess (so XL-BOMD and restarts have something to carry):
XL-BOMD's auxiliary density against its defining recursion and the published
coefficient properties, BAOAB against its splitting, bit-for-bit continuation
through a checkpoint file (aimd/checkpoint.py, aimd/integrators.py:125-163),
and the checkpoint's fingerprint check.
"""

import numpy as np
import pytest

from aimd_amd.backends.base import GradientResult


class ToyBackend:
    """Two Morse-like atoms; the 'density' is a 2x2 matrix relaxed towards f(x) from the
    guess by a fixed number of damped steps, and the energy carries the unrelaxed rest
    quadratically (an SCF stopped early, as XL-BOMD assumes)."""
    name = "toy"
    supports_density_guess = True
    reuse_density = True
    label = "TOY"
    nao = 2
    density_shape = (2, 2)
    basis = "toy-basis"

    def __init__(self, cycles=3):
        self.cycles = cycles
        self._guess = None
        self._last = None
        self.calls = 0

    def set_density_guess(self, d):
        self._guess = np.array(d, dtype=float)

    def _target(self, x):
        r = np.linalg.norm(x[1] - x[0])
        return np.array([[np.exp(-r), 0.1 * r], [0.1 * r, 1.0 - np.exp(-r)]])

    def compute(self, x):
        self.calls += 1
        x = np.asarray(x, dtype=float)
        f = self._target(x)
        d = self._guess if self._guess is not None else (self._last if self._last is not None else np.eye(2))
        self._guess = None
        for _ in range(self.cycles):
            d = f + 0.3 * (d - f)
        self._last = d.copy()
        dvec = x[1] - x[0]
        r = np.linalg.norm(dvec)
        e = (1.0 - np.exp(-(r - 2.0))) ** 2 + 0.5 * np.sum((d - f) ** 2)
        dedr = 2.0 * (1.0 - np.exp(-(r - 2.0))) * np.exp(-(r - 2.0))
        g = np.zeros_like(x)
        g[1] = dedr * dvec / r
        g[0] = -g[1]
        return GradientResult(energy=e, gradient=g, converged=True, density=d.copy(), dipole=np.zeros(3),
                              info={"scf_iterations": self.cycles})


def _system():
    from aimd_amd.md import System
    return System.create(["H", "H"], [[0.0, 0.0, 0.0], [0.0, 0.0, 2.3]], [[0.0, 0.0, 1e-3], [0.0, 0.0, -1e-3]])


def test_xl_coefficients_are_dissipative_moment_free():
    """sum c_k = sum k c_k = 0 (constant and linear densities untouched) and the odd
    moments vanish below order 2K - 3 (the property the 2009 table is built on)."""
    from aimd_amd.md import XL_TABLE
    for k, (_, _, c) in XL_TABLE.items():
        c = np.array(c, dtype=float)
        kk = np.arange(len(c), dtype=float)
        assert c.sum() == 0 and (kk * c).sum() == 0
        for j in range(3, 2 * k - 3, 2):
            assert (kk ** j * c).sum() == 0, (k, j)


def test_auxiliary_density_recursion_numpy_and_torch():
    import torch
    from aimd_amd.md import AuxiliaryDensity, XL_TABLE
    rng = np.random.default_rng(0)
    aux_n, aux_t = AuxiliaryDensity(5), AuxiliaryDensity(5)
    d0 = rng.standard_normal((3, 3))
    aux_n.reset(d0)
    aux_t.reset(torch.tensor(d0))
    hist = [d0.copy() for _ in range(6)]
    kappa, alpha, c = XL_TABLE[5]
    for step in range(7):
        d = rng.standard_normal((3, 3))
        want = 2 * hist[0] - hist[1] + kappa * (d - hist[0]) + alpha * sum(ck * h for ck, h in zip(c, hist))
        hist = [want] + hist[:-1]
        got_n = aux_n.propagate(d)
        got_t = aux_t.propagate(torch.tensor(d)).numpy()
        np.testing.assert_allclose(got_n, want, rtol=0, atol=1e-12)
        np.testing.assert_allclose(got_t, want, rtol=0, atol=1e-12)
    # a constant density is a fixed point once the history holds it
    aux = AuxiliaryDensity(5)
    aux.reset(d0)
    np.testing.assert_allclose(aux.propagate(d0), d0, rtol=0, atol=1e-14)
    with pytest.raises(ValueError, match="order K"):
        AuxiliaryDensity(2)


def test_xlbomd_feeds_the_propagated_guess():
    from aimd_amd.md import XLBOMD
    be = ToyBackend()
    sysm = _system()
    xl = XLBOMD(be, 0.5, k=5)
    xl.initialize(sysm)
    seen = []
    orig = be.set_density_guess

    def spy(d):
        seen.append(np.array(d))
        orig(d)

    be.set_density_guess = spy
    for _ in range(4):
        xl.step(sysm)
    assert len(seen) == 4
    np.testing.assert_array_equal(seen[-1], xl.aux.current)


def test_baoab_splitting_and_heat():
    from aimd_amd.md import LangevinBAOAB
    be = ToyBackend()
    sysm = _system()
    lb = LangevinBAOAB(be, 0.5, 300.0, friction_per_fs=0.01, rng=4)
    lb.initialize(sysm)
    x0, v0 = sysm.positions.copy(), sysm.velocities.copy()
    a0 = -be.compute(x0).gradient / sysm.masses[:, None]
    rng = np.random.default_rng(4)
    dt = lb.dt
    v = v0 + 0.5 * dt * a0
    x = x0 + 0.5 * dt * v
    ek0 = 0.5 * np.sum(sysm.masses[:, None] * v * v)
    from aimd_amd.units import KB_AU
    v = lb.c1 * v + lb.c2 * np.sqrt(KB_AU * 300.0 / sysm.masses)[:, None] * rng.normal(size=v.shape)
    ek1 = 0.5 * np.sum(sysm.masses[:, None] * v * v)
    x = x + 0.5 * dt * v
    lb.step(sysm)
    np.testing.assert_allclose(sysm.positions, x, rtol=0, atol=1e-15)
    assert lb.heat == pytest.approx(ek1 - ek0, rel=1e-12)
    assert not sysm.com_removed


@pytest.mark.parametrize("kind", ["VelocityVerlet", "XLBOMD", "LangevinBAOAB"])
def test_checkpoint_restart_is_bit_for_bit(tmp_path, kind):
    from aimd_amd import md
    n, m = 5, 6

    def make(be):
        if kind == "LangevinBAOAB":
            return md.LangevinBAOAB(be, 0.5, 300.0, friction_per_fs=0.01, rng=7)
        if kind == "XLBOMD":
            return md.XLBOMD(be, 0.5, k=5)
        return md.VelocityVerlet(be, 0.5)

    ref_sys = _system()
    ref = make(ToyBackend())
    ref.initialize(ref_sys)
    for _ in range(n + m):
        ref.step(ref_sys)
    s1 = _system()
    run = make(ToyBackend())
    run.initialize(s1)
    for _ in range(n):
        run.step(s1)
    path = md.save_checkpoint(tmp_path / "run.npz", s1, run, n)
    ck = md.load_checkpoint(path)
    assert ck.step == n and ck.backend_info["name"] == "toy" and ck.backend_info["density_shape"] == [2, 2]
    be2 = ToyBackend()
    ck.check_backend(be2)
    s2 = _system()
    cont = make(be2)
    ck.restore(s2, cont)
    for _ in range(m):
        cont.step(s2)
    np.testing.assert_array_equal(s2.positions, ref_sys.positions)
    np.testing.assert_array_equal(s2.velocities, ref_sys.velocities)
    assert cont.result.energy == ref.result.energy
    assert cont.conserved_energy(s2) == ref.conserved_energy(ref_sys)


def test_checkpoint_rejects_other_backends_and_damaged_files(tmp_path):
    from aimd_amd import md
    be = ToyBackend()
    s = _system()
    vv = md.VelocityVerlet(be, 0.5)
    vv.initialize(s)
    p = md.save_checkpoint(tmp_path / "c.npz", s, vv, 0)
    other = ToyBackend()
    other.basis = "another"
    with pytest.raises(ValueError, match="basis"):
        md.load_checkpoint(p).check_backend(other)
    bad = tmp_path / "bad.npz"
    bad.write_bytes(p.read_bytes()[:100])
    with pytest.raises(ValueError):
        md.load_checkpoint(bad)
    with pytest.raises(ValueError, match="cannot load"):
        md.XLBOMD(ToyBackend(), 0.5).load_state_dict(vv.state_dict())
