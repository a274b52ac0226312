"""The NumPy integrator that checks the device MD step, pinned against what
the mathematics fixes: the constraint stages reach their tolerances, the
thermostatted velocity update is the exact solution of dv/dt = a - vxi v
over dt/2, the Yoshida-Suzuki weights satisfy the order conditions, and the
Nose-Hoover chain step conserves its extended energy and thermalises a set
of harmonic oscillators."""
import math

import numpy as np
import pytest

from thrid_amd import md
from helpers import water648
import md_reference as mr


@pytest.mark.parametrize("vxi", [-3e-2, -1e-4, 0.0, 2e-6, 1e-3, 5e-2])
def test_kick_factors_closed_form(vxi):
    """v(dt/2) = v exp(-vxi dt/2) + a (1 - exp(-vxi dt/2)) / vxi."""
    dt = 2.0
    nhc = mr.Nhc(3, 300.0, 0.01, 100, 1, 1)
    nhc.v[0] = vxi
    s2, bbv = nhc.kick_factors(dt)
    h = 0.5 * dt
    assert s2 == pytest.approx(math.exp(-vxi * h), rel=1e-15)
    exact = h if vxi == 0.0 else -math.expm1(-vxi * h) / vxi
    assert bbv == pytest.approx(exact, rel=1e-13)


@pytest.mark.parametrize("nyosh", [1, 3, 5])
def test_yoshida_suzuki_order_conditions(nyosh):
    w = np.array(mr.ys_coeff(nyosh))
    assert len(w) == nyosh
    assert abs(w.sum() - 1.0) < 1e-15
    assert np.allclose(w, w[::-1], rtol=0, atol=0)       # symmetric
    if nyosh > 1:
        assert abs((w ** 3).sum()) < 1e-15                # fourth order


def test_shake_and_rattle_reach_tolerance():
    s = water648()
    ij, d, gs = md.rigid_constraints(s)
    minv = 1.0 / np.asarray(s["mass"], np.float64)
    d2 = d ** 2
    groups = mr._groups(ij, gs)
    rng = np.random.default_rng(3)
    old = np.array(s["pos"], np.float64)
    vel = rng.normal(scale=0.01, size=old.shape)
    pos = old + 1.0 * vel
    mr.shake(pos, vel, old, minv, ij, d2, groups, 1.0, eps=1e-12)
    r2 = ((pos[ij[:, 0] - 1] - pos[ij[:, 1] - 1]) ** 2).sum(1)
    assert np.max(np.abs(r2 - d2) / d2) < 1e-12
    mr.rattle_v(pos, vel, minv, ij, d2, groups, eps=1e-12)
    r = pos[ij[:, 0] - 1] - pos[ij[:, 1] - 1]
    v = vel[ij[:, 0] - 1] - vel[ij[:, 1] - 1]
    # rattlep_v's stopping test, (r.v)^2 / r^2 * v^2 < eps^2, holds for every
    # constraint at once (not only for the last one each group visited)
    rv = (r * v).sum(1)
    assert np.max(rv * rv / (r * r).sum(1) * (v * v).sum(1)) < 1e-24
    # momentum is untouched by both stages
    m = 1.0 / minv
    assert np.allclose((m[:, None] * vel).sum(0), (m[:, None] * (pos - old)).sum(0), atol=1e-9)


def _oscillators(n=60, seed=5):
    rng = np.random.default_rng(seed)
    mass = rng.uniform(1.0, 16.0, n)
    omega = rng.uniform(0.03, 0.09, n)              # 1/fs
    k = omega ** 2 * mass / md.MD_CONV               # eref / A^2
    return mass, k


def test_nhc_conserves_extended_energy_and_thermalises():
    mass, k = _oscillators()
    n = len(mass)
    minv = 1.0 / mass
    rng = np.random.default_rng(7)
    t0, target = 700.0, 300.0
    # start at equipartition at t0 (positions and velocities)
    vel = rng.normal(size=(n, 3)) * np.sqrt(md.KB_EREF * t0 * md.MD_CONV / mass)[:, None]
    pos = rng.normal(size=(n, 3)) * np.sqrt(md.KB_EREF * t0 / k)[:, None]
    nhc = mr.Nhc(4, target, 0.02, 3 * n, 3, 1)
    rows = []

    def force(p):
        return -k[:, None] * p, 0.5 * (k[:, None] * p * p).sum()

    def record(p, v, pe):
        ke = 0.5 * (mass[:, None] * v * v).sum() / md.MD_CONV
        rows.append((ke, ke + pe + nhc.conserved_extra()))

    mr.nhc_steps(force, pos, vel, minv, 1.0, 6000, nhc, record=record)
    rows = np.array(rows)
    ke, h = rows[:, 0], rows[:, 1]
    temp = 2.0 * ke / (3 * n * md.KB_EREF)
    # the extended energy stays put while the oscillators lose half of theirs
    assert np.std(h) < 2e-3 * np.mean(ke)
    assert abs(h[-500:].mean() - h[:500].mean()) < 1e-3 * np.mean(ke)
    assert temp[:20].mean() > 450.0
    assert abs(temp[3000:].mean() - target) < 0.08 * target


def test_respa_with_one_inner_step_is_velocity_verlet(oracle_mod):
    """With every term on the long level and one inner step, the RESPA
    restatement performs the same operations as the plain step (the short
    forces are zero, and adding them changes nothing)."""
    from thrid_amd import abi
    s = water648()
    ij, d, gs = md.rigid_constraints(s)
    v0 = md.maxwell_boltzmann(s, 300.0, seed=2)
    mts = abi.mts_all(abi.MTS_LONG)
    p1, v1, r1 = mr.run(oracle_mod.Oracle(s), s, v0, ij, d, gs, 1.0, 3)
    p2, v2, r2 = mr.run_respa(oracle_mod.Oracle(s), s, v0, ij, d, gs, 1.0, 1, 3, mts)
    np.testing.assert_array_equal(p1, p2)
    np.testing.assert_array_equal(v1, v2)
    np.testing.assert_array_equal(r1, r2)
