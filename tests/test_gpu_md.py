"""The device-resident NVE step (velocity Verlet + RATTLE) against the NumPy
integrator driven by the oracle, and its conservation properties."""
import numpy as np
import pytest

from thrid_amd import md
from helpers import water648
import md_reference

pytestmark = pytest.mark.gpu


def _ctx(s, prec, dt):
    from thrid_amd.pgk import ForceContext
    c = ForceContext(s, prec)
    ij, d, gs = md.rigid_constraints(s)
    c.md_setup(s["mass"], ij, d, gs, dt)
    return c, (ij, d, gs)


def test_md_trajectory_matches_reference(oracle_mod):
    s = water648()
    v0 = md.maxwell_boltzmann(s, 300.0, seed=2)
    c, (ij, d, gs) = _ctx(s, "fp64", 1.0)
    c.md_set_velocities(v0)
    rows = c.md_run(5, 20, s["rcut_book"], 1)
    p, v = c.md_state()
    rp, rv, rrows = md_reference.run(oracle_mod.Oracle(s), s, v0, ij, d, gs, 1.0, 5)
    assert np.max(np.abs(p - rp)) < 1e-9
    assert np.max(np.abs(v - rv)) < 1e-9 * np.max(np.abs(rv))
    np.testing.assert_allclose(rows, rrows, rtol=1e-9, atol=1e-9)


def test_md_general_groups_match_reference(oracle_mod):
    """Every second water keeps only its two O-H constraints: those groups
    are not triangles, so the step takes the separate kick/drift and
    constraint kernels with the global-memory group path."""
    s = water648()
    ij, d, gs = md.rigid_constraints(s)
    keep, starts = [], [0]
    for g in range(len(gs) - 1):
        ks = list(range(gs[g], gs[g + 1]))
        if g % 2 == 1:
            ks = [k for k in ks if s["atype"][ij[k, 0] - 1] != s["atype"][ij[k, 1] - 1]]
        keep += ks
        starts.append(len(keep))
    ij2, d2, gs2 = ij[keep], d[keep], np.asarray(starts, np.int32)
    assert len(ij2) < len(ij)
    v0 = md.maxwell_boltzmann(s, 300.0, seed=4)
    from thrid_amd.pgk import ForceContext
    c = ForceContext(s, "fp64")
    c.md_setup(s["mass"], ij2, d2, gs2, 1.0)
    c.md_set_velocities(v0)
    rows = c.md_run(5, 20, s["rcut_book"], 1)
    p, v = c.md_state()
    rp, rv, rrows = md_reference.run(oracle_mod.Oracle(s), s, v0, ij2, d2, gs2, 1.0, 5)
    assert np.max(np.abs(p - rp)) < 1e-9
    assert np.max(np.abs(v - rv)) < 1e-9 * np.max(np.abs(rv))
    np.testing.assert_allclose(rows, rrows, rtol=1e-9, atol=1e-9)


@pytest.mark.parametrize("prec", ["fp64", "mixed"])
def test_md_graph_replay_matches_eager(prec):
    """Plain steps replayed from a captured graph (between list rebuilds,
    samples and host synchronisations) match launching every kernel eagerly:
    bit-identical in the mixed form (its sums are order-independent), to
    1e-12 in FP64 (whose force atomics add in any order)."""
    s = water648()
    v0 = md.maxwell_boltzmann(s, 300.0, seed=6)
    out = []
    for graphs in (False, True):
        c, _cons = _ctx(s, prec, 1.0)
        c.md_use_graphs(single=graphs)
        c.md_set_velocities(v0)
        rows = c.md_run(47, 10, s["rcut_book"], 7)
        rows2 = c.md_run(13, 5, s["rcut_book"], 13)
        p, v = c.md_state()
        out.append((rows, rows2, p, v))
        c.close()
    for a, b in zip(out[0], out[1]):
        if prec == "mixed":
            assert np.array_equal(a, b)
        else:
            assert np.max(np.abs(a - b)) <= 1e-12 * max(1.0, np.max(np.abs(a)))


@pytest.mark.parametrize("prec", ["fp64", "mixed"])
def test_md_nve_conserves_energy_and_constraints(prec):
    s = water648()
    c, (ij, d, gs) = _ctx(s, prec, 1.0)
    c.md_set_velocities(md.maxwell_boltzmann(s, 300.0, seed=3))
    rows = c.md_run(1000, 20, s["rcut_book"], 10)
    e, ke = rows[:, 2], rows[:, 0]
    print(prec, "E fluctuation / KE", np.std(e) / np.mean(ke),
          "drift eref/ns/atom", md.drift_per_ns(rows, 10, 1.0, s["natom"]))
    assert np.std(e) < 5e-3 * np.mean(ke)
    assert abs(e[-1] - e[0]) < 1e-2 * np.mean(ke)
    p, _v = c.md_state()
    r = np.linalg.norm(p[ij[:, 0] - 1] - p[ij[:, 1] - 1], axis=1)
    assert np.max(np.abs(r - d)) < 1e-8


def test_md_batch_matches_single_runs():
    """Replicas stepped together (pgk_md_run_batch, interleaved streams, with
    and without per-replica graph replay) give each replica the trajectory it
    gets alone: bit-identical in the mixed form.  Three uneven replicas:
    water, n-butane (no constraints) and a water/propanol mixture, the first
    and last under a Nose-Hoover chain."""
    from thrid_amd import systems
    from thrid_amd.pgk import ForceContext, md_run_batch
    reps = [systems.mixture_box(1500, "water", seed=11, name="w"),
            systems.mixture_box(1800, "butane", seed=12, name="b"),
            systems.mixture_box(2100, "mix", seed=13, alcohol_frac=0.1, name="m")]
    for s in reps:   # relax the generated starts (huge initial forces)
        c = ForceContext(s, "fp64")
        s["pos"] = md.relax(c, s, md.rigid_constraints(s))
        c.close()

    def make(s, seed):
        c = ForceContext(s, "mixed")
        ij, d, gs = md.rigid_constraints(s)
        c.md_setup(s["mass"], ij, d, gs, 1.0)
        c.md_set_velocities(md.maxwell_boltzmann(s, 300.0, seed=seed))
        if seed % 2 == 0:   # the first and last replica under a thermostat
            c.md_nhc(320.0, 0.01, 3 * s["natom"] - len(d) - 3)
        return c

    single = []
    for k, s in enumerate(reps):
        c = make(s, 20 + k)
        rows = c.md_run(25, 10, s["rcut_book"], 5)
        single.append((rows,) + c.md_state())
        c.close()
    for graphs in (False, True):   # replicas launched eagerly, then from graphs
        ctxs = [make(s, 20 + k) for k, s in enumerate(reps)]
        for c in ctxs:
            c.md_use_graphs(single=False, batch=graphs)
        rows = md_run_batch(ctxs, 25, 10, [s["rcut_book"] for s in reps], 5)
        for k, c in enumerate(ctxs):
            p, v = c.md_state()
            assert np.array_equal(rows[k], single[k][0])
            assert np.array_equal(p, single[k][1])
            assert np.array_equal(v, single[k][2])
            c.close()


def _nfree(s, d):
    return 3 * s["natom"] - len(d) - 3


def test_md_nhc_matches_reference(oracle_mod):
    """The NVT step (Nose-Hoover chain of 4, Yoshida-Suzuki order 3, two
    nhcint sub-steps) against the NumPy integrator on the oracle's forces:
    positions, velocities, the chain's state and the energy samples (E the
    conserved quantity)."""
    s = water648()
    v0 = md.maxwell_boltzmann(s, 300.0, seed=8)
    c, (ij, d, gs) = _ctx(s, "fp64", 1.0)
    c.md_set_velocities(v0)
    c.md_nhc(250.0, 0.01, _nfree(s, d), mchain=4, nyosh=3, next=2)
    rows = c.md_run(5, 20, s["rcut_book"], 1)
    p, v = c.md_state()
    x, vx = c.md_nhc_state()
    nhc = md_reference.Nhc(4, 250.0, 0.01, _nfree(s, d), 3, 2)
    mass = np.asarray(s["mass"], np.float64)
    rrows = []

    def record(pos, vel, pe):
        ke = 0.5 * (mass[:, None] * vel * vel).sum() / md.MD_CONV
        rrows.append([ke, pe, ke + pe + nhc.conserved_extra()])

    rp, rv = md_reference.run_nhc(oracle_mod.Oracle(s), s, v0, ij, d, gs, 1.0, 5, nhc,
                                  record=record)
    assert np.max(np.abs(p - rp)) < 1e-9
    assert np.max(np.abs(v - rv)) < 1e-9 * np.max(np.abs(rv))
    assert np.max(np.abs(vx - nhc.v)) < 1e-9 * np.max(np.abs(nhc.v))
    assert np.max(np.abs(x - nhc.x)) < 1e-9 * np.max(np.abs(nhc.x))
    assert np.max(np.abs(nhc.v)) > 1e-4      # the chain did move
    np.testing.assert_allclose(rows, np.array(rrows), rtol=1e-9, atol=1e-9)


@pytest.mark.parametrize("prec", ["fp64", "mixed"])
def test_md_nhc_thermalises_and_conserves(prec):
    """2000 NVT steps of the relaxed box from 300 K towards 360 K: the
    extended energy stays put, the temperature settles at the target and the
    constraints hold."""
    from thrid_amd.pgk import ForceContext
    s = water648()
    c = ForceContext(s, "fp64")
    s["pos"] = md.relax(c, s, md.rigid_constraints(s))
    c.close()
    c, (ij, d, gs) = _ctx(s, prec, 1.0)
    c.md_set_velocities(md.maxwell_boltzmann(s, 300.0, seed=9))
    c.md_nhc(360.0, 0.02, _nfree(s, d))
    rows = c.md_run(2000, 20, s["rcut_book"], 10)
    ke, h = rows[:, 0], rows[:, 2]
    temp = 2.0 * ke / (_nfree(s, d) * md.KB_EREF)
    print(prec, "H fluctuation / KE", np.std(h) / np.mean(ke), "T by 200 fs",
          [round(temp[k:k + 20].mean(), 1) for k in range(0, 200, 20)])
    assert np.std(h) < 5e-3 * np.mean(ke)
    assert abs(h[-1] - h[0]) < 1e-2 * np.mean(ke)
    assert abs(temp[100:].mean() - 360.0) < 0.05 * 360.0
    p, _v = c.md_state()
    r = np.linalg.norm(p[ij[:, 0] - 1] - p[ij[:, 1] - 1], axis=1)
    assert np.max(np.abs(r - d)) < 1e-8



def _calcom(s, pos):
    """calcom (calcom.F90:1-147): sum m x over each molecule's atoms in index
    order times 1/M (a one-atom molecule: its position), wrapped once."""
    mass = np.asarray(s["mass"], np.float64)
    molid = np.asarray(s["molid"])
    box = np.asarray(s["box"], np.float64)
    com = np.zeros((int(molid.max()), 3))
    for m in range(com.shape[0]):
        atoms = np.nonzero(molid == m + 1)[0]
        if len(atoms) == 1:
            c = pos[atoms[0]].copy()
        else:
            mm, c = 0.0, np.zeros(3)
            for i in atoms:
                mm = mm + mass[i]
                c = c + mass[i] * pos[i]
            c = c * (1.0 / mm)
        c = np.where(c < 0.0, c + box, np.where(c >= box, c - box, c))
        com[m] = c
    return com


@pytest.mark.parametrize("prec", ["fp64", "mixed"])
@pytest.mark.parametrize("mode", ["bulk", "volume", "volume_mole"])
def test_md_heatf_series_matches_fused_call(prec, mode):
    """The heat-flux rows of the MD loop (steps 0, 5, 10) against the fused
    calforce_heatf call on the same state: the start, and the end of the run.
    volume_mole: every water by its molecule's centre of mass, which the MD
    loop computes on the device and the fused call takes from the host."""
    from thrid_amd import abi
    s = water648()
    c, _cons = _ctx(s, prec, 1.0)
    c.md_set_velocities(md.maxwell_boltzmann(s, 300.0, seed=12))
    p0, v0 = c.md_state()
    lz = s["box"][2]
    if mode == "bulk":
        hf = dict(mode=abi.HF_BULK, z1=np.array([0.0]), z2=np.array([lz]))
    else:
        z = np.linspace(0.0, lz, 5)
        hf = dict(mode=abi.HF_VOLUME, z1=z[:-1], z2=z[1:])
        if mode == "volume_mole":
            hf["hftyp"] = np.full(s["natom"], abi.HFTYP_MOLE, np.int32)
    c.md_heatf(hf, 5)
    rows = c.md_run(10, 20, s["rcut_book"], 5)
    series = c.md_heatf_series()
    assert series.shape == (3, len(hf["z1"]), abi.NHFJ, 3)
    p1, v1 = c.md_state()
    ene = c.energies()
    # sampling the heat flux leaves the trajectory and the energy records alone
    c2, _ = _ctx(s, prec, 1.0)
    c2.md_set_velocities(md.maxwell_boltzmann(s, 300.0, seed=12))
    rows2 = c2.md_run(10, 20, s["rcut_book"], 5)
    p2, v2 = c2.md_state()
    ene2 = c2.energies()
    c2.close()
    tol = 0.0 if prec == "mixed" else 1e-12
    for a, b in ((rows, rows2), (p1, p2), (v1, v2), (ene[0], ene2[0])):
        assert np.max(np.abs(a - b)) <= tol * max(1.0, np.max(np.abs(b)))
    m = s["mass"] / md.MD_CONV          # m v^2 / 2 in eref, as the MD loop uses
    if mode == "volume_mole":
        hf["molecom"] = _calcom(s, p1)
    end = c.calforce_heatf(hf, v1, m)["hf"]
    c.set_positions(p0)
    if mode == "volume_mole":
        hf["molecom"] = _calcom(s, p0)
    start = c.calforce_heatf(hf, v0)["hf"]
    for got, ref in ((series[2], end), (series[0], start)):
        scale = np.max(np.abs(ref))
        assert scale > 0
        assert np.max(np.abs(got - ref)) <= 1e-11 * scale
    assert np.max(np.abs(series[2] - series[0])) > 1e-3 * np.max(np.abs(series[0]))
    c.md_heatf(None, 0)
    c.md_run(2, 20, s["rcut_book"], 1)
    assert c.md_heatf_series().shape[0] == 0


def _respa_system():
    from helpers import small_slab
    from thrid_amd import abi
    s = small_slab("rb")
    # bonded terms on the short level, non-bonded and reciprocal on the long
    mts = abi.Mts(abi.MTS_SHORT, abi.MTS_SHORT, abi.MTS_SHORT, abi.MTS_SHORT,
                  abi.MTS_LONG, abi.MTS_LONG, abi.MTS_LONG)
    return s, mts


def test_md_respa_matches_reference(oracle_mod):
    """3 outer steps of RESPA (3 inner steps, bonded terms short) on the
    water/propanol slab against the NumPy RESPA on the oracle's forces."""
    from thrid_amd.pgk import ForceContext
    s, mts = _respa_system()
    ij, d, gs = md.rigid_constraints(s)
    v0 = md.maxwell_boltzmann(s, 300.0, seed=14)
    c = ForceContext(s, "fp64")
    c.md_setup(s["mass"], ij, d, gs, 1.5)
    c.md_respa(3, mts)
    c.md_set_velocities(v0)
    rows = c.md_run(3, 20, s["rcut_book"], 1)
    p, v = c.md_state()
    c.close()
    rp, rv, rrows = md_reference.run_respa(oracle_mod.Oracle(s), s, v0, ij, d, gs, 1.5, 3, 3, mts)
    assert np.max(np.abs(p - rp)) < 1e-9
    assert np.max(np.abs(v - rv)) < 1e-9 * np.max(np.abs(rv))
    np.testing.assert_allclose(rows, rrows, rtol=1e-9, atol=1e-9)


@pytest.mark.parametrize("prec", ["fp64", "mixed"])
def test_md_respa_conserves_energy(prec):
    """300 outer steps of 2 fs with 4 inner steps of 0.5 fs: the total energy
    stays put and the rigid waters stay rigid."""
    from thrid_amd.pgk import ForceContext
    s, mts = _respa_system()
    cons = md.rigid_constraints(s)
    c = ForceContext(s, "fp64")
    s = dict(s)
    s["pos"] = md.relax(c, s, cons)
    c.close()
    c = ForceContext(s, prec)
    c.md_setup(s["mass"], *cons, 2.0)
    c.md_respa(4, mts)
    c.md_set_velocities(md.maxwell_boltzmann(s, 300.0, seed=15))
    rows = c.md_run(300, 10, s["rcut_book"], 5)
    e, ke = rows[:, 2], rows[:, 0]
    print(prec, "RESPA E fluctuation / KE", np.std(e) / np.mean(ke),
          "end - start", (e[-1] - e[0]) / np.mean(ke))
    assert np.std(e) < 1e-2 * np.mean(ke)
    assert abs(e[-1] - e[0]) < 2e-2 * np.mean(ke)
    p, _v = c.md_state()
    ij, d, _gs = cons
    r = np.linalg.norm(p[ij[:, 0] - 1] - p[ij[:, 1] - 1], axis=1)
    assert np.max(np.abs(r - d)) < 1e-8
