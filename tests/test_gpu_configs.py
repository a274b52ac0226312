"""GPU parity at the workloads' own sizes, against the oracle.

- water/1-propanol slab: 4096 waters + 512 propanols (18,432 atoms), 50 x 50 x
  150 A, SPME 64 x 64 x 192 order 4, 12 A cutoff: calforce in FP64 (rms force
  error <= 1e-5) and mixed (<= 1e-4), and calforce_hf's volume-mode per-atom
  arrays over 10 slabs (FP64, 1e-10 of the largest entry);
- 100,001-atom water + Na+/Cl- box, 100.1 A, SPME 128^3, 10 A cutoff:
  calforce in FP64 (<= 1e-5) and mixed (<= 1e-4), energies and virials;
- 64 uneven replicas (water, n-butane, water/propanol; 1,500-12,000 atoms):
  every replica's step-0 forces in the mixed form (<= 1e-4), and a few MD
  steps stepped together bit-identical to each replica run alone.

The oracle is the C restatement of the reference's FP64 path; it takes tens of
seconds at these sizes, hence the longer per-test time limits.
"""
import multiprocessing as mp

import numpy as np
import pytest

from thrid_amd import abi, md, systems
from helpers import rms_rel, ene

pytestmark = pytest.mark.gpu


def _ctx(s, prec):
    from thrid_amd.pgk import ForceContext
    c = ForceContext(s, prec)
    c.mklist(s["rcut_book"])
    return c


def _energy_checks(got, ref, tol, names):
    for name in names:
        g, r = ene(got, name), ene(ref, name)
        assert abs(g - r) <= tol * max(1.0, abs(r)), f"{name}: {g!r} vs {r!r}"


@pytest.fixture(scope="module")
def slab_full(oracle_mod):
    s = systems.propanol_water_slab()
    assert s["natom"] == 4096 * 3 + 512 * 12
    o = oracle_mod.Oracle(s)
    o.mklist(s["rcut_book"], 1)
    return s, o, o.calforce(abi.MTS_LONG)


@pytest.mark.timeout(600)
@pytest.mark.parametrize("prec,tol", [("fp64", 1e-5), ("mixed", 1e-4)])
def test_slab_full_calforce(slab_full, prec, tol):
    s, _o, ref = slab_full
    c = _ctx(s, prec)
    got = c.calforce(abi.MTS_LONG)
    err = rms_rel(got["force"], ref["force"])
    print(f"slab 18432 {prec}: rms rel force err {err:.3e}")
    assert err <= tol
    etol = 1e-9 if prec == "fp64" else 1e-4
    _energy_checks(got, ref, etol, ["pot_elc", "pot_vdw", "pot_ewk", "pot_bond", "pot_angl",
                                    "pot_torsrb", "pot_elc14", "pot_vdw14", "pot_tot"])
    if prec == "fp64":
        vs = max(1.0, float(np.max(np.abs(ref["virt"]))))
        assert float(np.max(np.abs(got["virt"] - ref["virt"]))) <= 1e-9 * vs
    c.close()


@pytest.mark.timeout(600)
def test_slab_full_calforce_hf_volume(slab_full):
    s, o, _ref = slab_full
    z = np.linspace(0.0, s["box"][2], 11)
    hf = dict(mode=abi.HF_VOLUME, z1=z[:-1], z2=z[1:])
    ref = o.calforce_hf(hf)
    c = _ctx(s, "fp64")
    got = c.calforce_hf(hf)
    assert rms_rel(got["force"], ref["force"]) < 1e-9
    pe = max(1.0, float(np.max(np.abs(ref["pot_atm"]))))
    assert float(np.max(np.abs(got["pot_atm"] - ref["pot_atm"]))) <= 1e-10 * pe
    w = max(1.0, float(np.max(np.abs(ref["viri_atm"]))))
    dw = float(np.max(np.abs(got["viri_atm"] - ref["viri_atm"])))
    print(f"slab 18432 calforce_hf: max |dW| {dw:.3e} of {w:.3e}")
    assert dw <= 1e-10 * w
    c.close()


@pytest.fixture(scope="module")
def box100k(oracle_mod):
    s = systems.water_ion_box_100k()
    o = oracle_mod.Oracle(s)
    o.mklist(s["rcut_book"], 1)
    return s, o.calforce(abi.MTS_LONG)


@pytest.mark.timeout(600)
@pytest.mark.parametrize("prec,tol", [("fp64", 1e-5), ("mixed", 1e-4), ("fp32", 1e-4)])
def test_100k_calforce(box100k, prec, tol):
    s, ref = box100k
    c = _ctx(s, prec)
    got = c.calforce(abi.MTS_LONG)
    err = rms_rel(got["force"], ref["force"])
    print(f"100k {prec}: rms rel force err {err:.3e}")
    assert err <= tol
    etol = 1e-9 if prec == "fp64" else 1e-4
    _energy_checks(got, ref, etol, ["pot_elc", "pot_vdw", "pot_ewk", "pot_ewc", "pot_tot"])
    vs = max(1.0, float(np.max(np.abs(ref["virt"]))))
    assert float(np.max(np.abs(got["virt"] - ref["virt"]))) <= (1e-9 if prec == "fp64" else 1e-3) * vs
    c.close()


@pytest.fixture(scope="module")
def ensemble64():
    # built in fresh interpreters: this process has already used the GPU, so
    # it must not fork
    return systems.replica_ensemble(64, workers=8, mp_context="spawn")


@pytest.mark.timeout(900)
def test_ensemble64_sizes_and_step0_forces(ensemble64, oracle_mod):
    sizes = np.array([s["natom"] for s in ensemble64])
    print("replica sizes", sizes.min(), sizes.max(), sizes.sum())
    assert len(ensemble64) == 64
    assert sizes.min() >= 1400 and sizes.max() <= 12100
    worst = 0.0
    for s in ensemble64:
        o = oracle_mod.Oracle(s)
        o.mklist(s["rcut_book"], 1)
        ref = o.calforce(abi.MTS_LONG)
        c = _ctx(s, "mixed")
        got = c.calforce(abi.MTS_LONG)
        c.close()
        err = rms_rel(got["force"], ref["force"])
        worst = max(worst, err)
        assert err <= 1e-4, (s["name"], err)
    print(f"64 replicas: worst rms rel force err (mixed) {worst:.3e}")


@pytest.mark.timeout(900)
def test_ensemble64_batch_bit_identical(ensemble64):
    from thrid_amd.pgk import ForceContext, md_run_batch
    reps = [dict(s) for s in ensemble64]
    cons = []
    quick = ((0.02, 15), (0.1, 15), (0.5, 15))
    for s in reps:
        cn = md.rigid_constraints(s)
        c = ForceContext(s, "fp64")
        s["pos"] = md.relax(c, s, cn, schedule=quick)
        c.close()
        cons.append(cn)

    def make(k):
        s = reps[k]
        c = ForceContext(s, "mixed")
        c.md_setup(s["mass"], *cons[k], 1.0)
        c.md_set_velocities(md.maxwell_boltzmann(s, 300.0, seed=500 + k))
        return c

    nstep = 6
    single = []
    for k in range(len(reps)):
        c = make(k)
        rows = c.md_run(nstep, 3, reps[k]["rcut_book"], 2)
        single.append((rows,) + c.md_state())
        c.close()
    ctxs = [make(k) for k in range(len(reps))]
    rows = md_run_batch(ctxs, nstep, 3, [s["rcut_book"] for s in reps], 2)
    for k, c in enumerate(ctxs):
        p, v = c.md_state()
        assert np.array_equal(p, single[k][1]), reps[k]["name"]
        assert np.array_equal(v, single[k][2]), reps[k]["name"]
        assert np.array_equal(rows[k], single[k][0]), reps[k]["name"]
        c.close()
