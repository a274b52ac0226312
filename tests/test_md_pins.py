"""The MD restatement (tests/md_reference.py), which checks the device MD step,
against the reference's own integrator routines compiled from source:
rattle_c (rattle.F90:6-272), rattlep_v (rattle_mtk.F90:278-...) and nhcint
(nhcint.F90:8-140), driven through oracle/ref_harness.F90 on the 648-atom water
box (216 rigid 3-site waters: three constraints per molecule).

Both sides evaluate the same IEEE double expressions in the same order, so the
positions, velocities and chain variables are required to agree to a few ulp
(observed: bit for bit).  The region Langevin thermostat is not pinned here:
the reference draws its random numbers from SFMT (SFMT.F90) while the library
uses SplitMix64, so the two agree in distribution only
(tests/test_langevin.py).
"""
import numpy as np
import pytest

from thrid_amd import md
from helpers import water648
import md_reference as mr


def _ulp_close(a, b, rel=4e-16):
    scale = max(1.0, float(np.max(np.abs(b))))
    return float(np.max(np.abs(np.asarray(a) - np.asarray(b)))) <= rel * scale * 8


@pytest.fixture()
def water_md(oracle_mod):
    if not oracle_mod.reference_available():
        pytest.skip("reference not built (oracle/build_ref.sh)")
    s = water648()
    ij, d, gs = md.rigid_constraints(s)
    ref = oracle_mod.Reference(s, use_mkexcl=True)
    return s, ref, ij, d, gs


def test_rattle_c_and_rattlep_v_match_reference(water_md):
    s, ref, ij, d, gs = water_md
    mass = np.asarray(s["mass"], np.float64)
    minv = 1.0 / mass
    d2 = d * d
    groups = mr._groups(ij, gs)
    p0 = np.array(s["pos"], np.float64)          # constraints hold at T
    v0 = md.maxwell_boltzmann(s, 600.0, seed=17)
    dt = 1.0
    p1 = p0 + v0 * dt                              # unconstrained drift
    # restatement
    pa, va = p1.copy(), v0.copy()
    mr.shake(pa, va, p0, minv, ij, d2, groups, dt, eps=1e-10, maxit=300)
    # the reference's rattle_c on the same state
    ref.set_positions(p1)
    ref.set_md(ij, d, gs, mass, v0)
    ref.rattle_c(p0, 1e-10, dt)
    pr, vr = ref.get_md()
    print("rattle_c: bit-identical positions", np.array_equal(pa, pr),
          "velocities", np.array_equal(va, vr))
    assert _ulp_close(pa, pr) and _ulp_close(va, vr)
    moved = np.max(np.abs(pa - p1))
    assert moved > 1e-3, "the drift broke no constraint: nothing was tested"
    # velocities at the constrained positions, then rattlep_v
    v2 = v0 + md.maxwell_boltzmann(s, 300.0, seed=18)
    vb = v2.copy()
    mr.rattle_v(pa, vb, minv, ij, d2, groups, eps=1e-10, maxit=300)
    ref.set_positions(pa)
    ref.set_md(ij, d, gs, mass, v2)
    ref.rattlep_v(1e-10, dt)
    _p, vr2 = ref.get_md()
    print("rattlep_v: bit-identical velocities", np.array_equal(vb, vr2))
    assert _ulp_close(vb, vr2)
    assert np.max(np.abs(vb - v2)) > 1e-4


@pytest.mark.parametrize("mchain,nyosh,nexti", [(4, 3, 2), (2, 5, 1), (3, 1, 3)])
def test_nhcint_matches_reference(water_md, mchain, nyosh, nexti):
    s, ref, ij, d, gs = water_md
    mass = np.asarray(s["mass"], np.float64)
    vel = md.maxwell_boltzmann(s, 420.0, seed=5)
    nfree = 3 * s["natom"] - len(d) - 3
    nhc = mr.Nhc(mchain, 300.0, 0.02, nfree, nyosh, nexti)
    nhc.x = [0.01 * (k + 1) for k in range(mchain)]
    nhc.v = [0.003 * (-1) ** k * (k + 1) for k in range(mchain)]
    x0, v0 = list(nhc.x), list(nhc.v)
    # the reference sums m v^2 over atoms in atom order, each as m (vx vx + vy vy + vz vz)
    kint = 0.0
    for i in range(s["natom"]):
        kint = kint + mass[i] * (vel[i, 0] * vel[i, 0] + vel[i, 1] * vel[i, 1]
                                 + vel[i, 2] * vel[i, 2])
    dt = 2.0
    for _ in range(3):
        nhc.half(kint, dt)
    ref.set_md(ij, d, gs, mass, vel)
    x, v = x0, v0
    for _ in range(3):
        x, v = ref.nhcint(mchain, nyosh, nexti, nfree, dt, nhc.gkt, nhc.ys, nhc.q, x, v)
    print("nhcint: bit-identical", np.array_equal(x, nhc.x), np.array_equal(v, nhc.v))
    assert _ulp_close(x, nhc.x) and _ulp_close(v, nhc.v)
    assert max(abs(a - b) for a, b in zip(v, v0)) > 1e-6
