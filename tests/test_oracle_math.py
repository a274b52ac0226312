"""The oracle against what the mathematics fixes, independent of the reference.

Each check is chosen so that a plausible slip in the restatement (a dropped
term, a wrong sign, factor or index, a transposed operand) fails it:
  * FFT against numpy.fft (sign convention, normalisation, axis order);
  * B-spline moduli against the closed form of Essmann et al. 1995, J. Chem.
    Phys. 103:8577, eq. 4.4 (fft_pme_init.F90:216-289);
  * Verlet list against an O(N^2) brute force with Fortran dnint
    (mklist.F90:8-234, mklist_cell.F90:8-352);
  * two-particle LJ and real-space Ewald energies/forces in closed form
    (ewaldr.F90:404-838);
  * the rock-salt Madelung constant through real + reciprocal + self energy
    (SPME, fft_pme.F90:544-690, fft_pme_init.F90:149-156);
  * F = -grad E by central differences for every term (pairs, SPME, bonded,
    1-4) on systems with bonded topology;
  * W_aa = -dE/d(eps_aa) under homogeneous scaling of box and coordinates
    (virial tensors of every class);
  * heat-flux sum rules: per-atom energies sum to the class energies and the
    per-atom virials summed over atoms (and over regions that tile the box)
    give the class virial (calvirihf.F90:8-405).
"""
import math

import numpy as np
import pytest

from thrid_amd import abi, systems
from helpers import ene, small_slab, water648


def bare_system(pos, charge, box, *, eps=0.0, rmin=3.0, alpha=0.35, rrcut=9.0, rcut=9.0,
                rcut_book=9.5, nfft=(32, 32, 32), order=6, nspltbl=None):
    """Point charges / one LJ type, no bonded terms, one molecule per atom."""
    n = len(pos)
    z2 = np.zeros((0, 2), np.int32)
    return dict(
        name="bare", natom=n, nmol=n, box=np.array(box, float),
        pos=np.ascontiguousarray(pos, np.float64), charge=np.asarray(charge, float),
        atype=np.ones(n, np.int32), molid=np.arange(1, n + 1, dtype=np.int32),
        mass=np.ones(n), lj_eps=np.array([[eps]]), lj_rmin=np.array([[rmin]]),
        inttyp=np.ones((1, 1), np.int32), alpha=alpha, rrcut=rrcut, rcut=rcut,
        rcut_book=rcut_book,
        nspltbl=int(nspltbl or round((rrcut + 2.0) / 0.01)), nfft=np.array(nfft, np.int32),
        pme_order=order,
        bond_ij=z2, bond_type=np.zeros(0, np.int32), bond_k=np.zeros(1), bond_r0=np.ones(1),
        angl_ijk=np.zeros((0, 3), np.int32), angl_type=np.zeros(0, np.int32),
        angl_k=np.zeros(1), angl_theta0=np.ones(1),
        tors_ijkl=np.zeros((0, 4), np.int32), tors_type=np.zeros(0, np.int32),
        tors_ring14=np.zeros(0), tors_divfac=np.ones(1, np.int32), tors_barhig=np.zeros(1),
        tors_phsang=np.zeros(1), tors_period=np.ones(1), tors_divfac_vdw=np.ones(1),
        tors_divfac_elc=np.ones(1), tors_if14=np.zeros(1, np.int32),
        torsrb_ijkl=np.zeros((0, 4), np.int32), torsrb_type=np.zeros(0, np.int32),
        torsrb_ring14=np.zeros(0), torsrb_c=np.zeros((1, 6)), torsrb_divfac_vdw=np.ones(1),
        torsrb_divfac_elc=np.ones(1), torsrb_if14=np.zeros(1, np.int32),
    )


# ---------------------------------------------------------------------------
@pytest.mark.parametrize("shape", [(8, 6, 5), (15, 12, 10), (16, 16, 16), (27, 9, 20)])
def test_fft_against_numpy(oracle_mod, shape):
    n1, n2, n3 = shape
    rng = np.random.default_rng(1)
    z = rng.normal(size=(n3, n2, n1)) + 1j * rng.normal(size=(n3, n2, n1))
    f = oracle_mod.fft3d(z, -1)
    np.testing.assert_allclose(f, np.fft.fftn(z), rtol=0, atol=1e-11 * np.abs(z).sum() ** 0.5)
    b = oracle_mod.fft3d(z, +1)
    np.testing.assert_allclose(b, np.fft.ifftn(z) * z.size, rtol=0,
                               atol=1e-11 * np.abs(z).sum() ** 0.5)


def _cardinal_bspline(p, u):
    """M_p(u) = 1/(p-1)! sum_k (-1)^k C(p,k) max(u-k, 0)^(p-1)."""
    s = 0.0
    for k in range(p + 1):
        s += (-1) ** k * math.comb(p, k) * max(u - k, 0.0) ** (p - 1)
    return s / math.factorial(p - 1)


@pytest.mark.parametrize("order,nfft", [(4, (24, 25, 30)), (5, (20, 27, 16)), (6, (32, 18, 45))])
def test_bspline_moduli_closed_form(oracle_mod, order, nfft):
    pos = np.array([[1.0, 1.0, 1.0], [3.0, 3.0, 3.0]])
    s = bare_system(pos, [1.0, -1.0], (20.0, 21.0, 22.0), nfft=nfft, order=order,
                    rrcut=8.0, rcut=8.0, rcut_book=8.5)
    o = oracle_mod.Oracle(s)
    _tbl, bsp, _rk = o.tables()
    off = 0
    for K in nfft:
        got = bsp[off:off + K]
        off += K
        m = np.arange(K)
        ks = np.arange(order - 1)
        mp = np.array([_cardinal_bspline(order, k + 1.0) for k in ks])
        den = np.abs((mp[None, :] * np.exp(2j * np.pi * m[:, None] * ks[None, :] / K)).sum(1)) ** 2
        ok = den > 1e-7
        np.testing.assert_allclose(got[ok], den[ok], rtol=1e-12)
        for k in np.nonzero(~ok)[0]:   # the tiny-value fix: mean of the neighbours
            assert got[k] == pytest.approx(0.5 * (got[k - 1] + got[(k + 1) % K]), rel=1e-12)


# ---------------------------------------------------------------------------
def _dnint(x):
    t = np.trunc(x)
    return t + np.sign(x) * (np.abs(x - t) >= 0.5)


def _brute_list(pos, box, rcb, excl_pairs):
    n = len(pos)
    idx = np.zeros(n, np.int32)
    lst = []
    # atoms 1..natom-1 own the pairs (mklist.F90:104), the last one no entry
    idx[n - 1] = -1
    for i in range(n - 1):
        d = pos[i] - pos[i + 1:]
        d = d - box * _dnint(d / box)
        r2 = d[:, 0] * d[:, 0] + d[:, 1] * d[:, 1] + d[:, 2] * d[:, 2]
        js = [i + 2 + k for k in np.nonzero(r2 < rcb * rcb)[0] if (i + 1, i + 2 + k) not in excl_pairs]
        idx[i] = len(lst) + 1
        lst += js if js else [0]
    idx[n - 1] = len(lst) + 1
    return idx, np.array(lst, np.int32)


@pytest.mark.parametrize("mode", [0, 1])
def test_list_brute_force(oracle_mod, mode):
    rng = np.random.default_rng(5)
    box = np.array([27.0, 28.5, 30.0])
    pos = rng.uniform(-1.0, 2.0, size=(420, 3)) * box   # images outside the box too
    s = bare_system(pos, np.zeros(420), box, rcut_book=9.0, rrcut=8.5, rcut=8.5)
    # a few exclusions, one atom with several
    ex = [(1, 5), (1, 9), (2, 3), (17, 400), (50, 51)]
    part = {}
    for i, j in ex:
        part.setdefault(i, []).append(j)
    eidx = np.zeros(421, np.int32)
    elst = []
    for a in range(1, 421):
        eidx[a - 1] = len(elst) + 1
        elst += sorted(part.get(a, [])) or [0]
    eidx[420] = len(elst) + 1
    o = oracle_mod.Oracle(s, excl=(eidx, np.array(elst, np.int32)))
    idx, lst = o.mklist(9.0, mode)
    bidx, blst = _brute_list(pos, box, 9.0, set(ex))
    assert len(blst) > 4000
    if mode == 0:
        assert np.array_equal(idx, bidx) and np.array_equal(lst, blst)
    else:   # the cell list visits neighbours in another order: compare per-atom sets
        ends = list(idx[1:]) + [len(lst) + 1]
        bends = list(bidx[1:]) + [len(blst) + 1]
        for i in range(420):
            a = sorted(lst[idx[i] - 1:ends[i] - 1])
            b = sorted(blst[bidx[i] - 1:bends[i] - 1])
            assert a == b, i
    if mode == 1:   # below -box the cell index has no cell for an atom
        p = pos.copy()
        p[7, 1] = -1.5 * box[1]
        o.set_positions(p)
        with pytest.raises(ValueError):
            o.mklist(9.0, 1)


# ---------------------------------------------------------------------------
def test_two_particle_closed_form(oracle_mod):
    from scipy.special import erfc
    box = (40.0, 41.0, 42.0)
    x1, x2 = np.array([10.0, 11.0, 12.0]), np.array([13.1, 12.2, 10.7])
    q1, q2 = 1.7, -0.9
    eps, rmin, alpha = 0.31, 3.4, 0.33
    s = bare_system(np.array([x1, x2]), [q1, q2], box, eps=eps, rmin=rmin, alpha=alpha,
                    rrcut=10.0, rcut=10.0, rcut_book=10.5, nfft=(40, 40, 40))
    o = oracle_mod.Oracle(s)
    o.mklist(10.5, 0)
    r = o.calforce(abi.MTS_LONG)
    d = x1 - x2
    rr = np.linalg.norm(d)
    e_lj = eps * ((rmin / rr) ** 12 - 2 * (rmin / rr) ** 6)
    f_lj = 12 * eps * ((rmin / rr) ** 12 - (rmin / rr) ** 6) / rr ** 2 * d
    assert ene(r, "pot_vdw") == pytest.approx(e_lj, rel=1e-12)
    np.testing.assert_allclose(r["for_viri_lj"][0], f_lj, rtol=1e-12)
    np.testing.assert_allclose(r["for_viri_lj"][1], -f_lj, rtol=1e-12)
    np.testing.assert_allclose(r["virt"][abi.TENS_NAMES.index("pot_virit_lj")],
                               np.outer(f_lj, d), rtol=1e-12, atol=1e-14)
    assert ene(r, "pot_viri_lj") == pytest.approx(float(f_lj @ d), rel=1e-12)
    e_re = q1 * q2 * erfc(alpha * rr) / rr
    assert ene(r, "pot_elc") == pytest.approx(e_re, rel=1e-8)
    self_e = -alpha / math.sqrt(math.pi) * (q1 * q1 + q2 * q2)
    assert ene(r, "pot_ewc") == pytest.approx(self_e, rel=1e-12)


def _nacl(ncell, a):
    pts = []
    for ix in range(2 * ncell):
        for iy in range(2 * ncell):
            for iz in range(2 * ncell):
                pts.append(((ix, iy, iz), 1.0 if (ix + iy + iz) % 2 == 0 else -1.0))
    pos = np.array([p for p, _ in pts], float) * (a / 2) + 0.13
    q = np.array([c for _, c in pts])
    return pos, q


def test_madelung_rock_salt(oracle_mod):
    a = 5.64
    pos, q = _nacl(3, a)                     # 216 ions, 16.92 A box
    L = 3 * a
    s = bare_system(pos, q, (L, L, L), alpha=0.5, rrcut=8.0, rcut=8.0, rcut_book=8.2,
                    nfft=(48, 48, 48), order=8)
    o = oracle_mod.Oracle(s)
    o.mklist(8.2, 0)
    r = o.calforce(abi.MTS_LONG)
    e = ene(r, "pot_elc") + ene(r, "pot_ewk") + ene(r, "pot_ewc")
    madelung = 1.747564594633182
    expect = -0.5 * len(q) * madelung / (a / 2)
    assert e == pytest.approx(expect, rel=2e-7)
    # a perfect lattice is in equilibrium
    assert np.max(np.abs(r["force"])) < 1e-5


# ---------------------------------------------------------------------------
def _systems_with_terms():
    return [("water648", water648()), ("slab_rb", small_slab("rb")),
            ("slab_periodic", small_slab("periodic"))]


@pytest.mark.parametrize("name", ["water648", "slab_rb", "slab_periodic"])
def test_force_is_minus_energy_gradient(oracle_mod, name):
    s = dict(_systems_with_terms())[name]
    o = oracle_mod.Oracle(s)
    idx, lst = o.mklist(s["rcut_book"], 0)
    f = o.calforce(abi.MTS_LONG)["force"]
    rng = np.random.default_rng(3)
    h = 1e-5
    for a in rng.choice(s["natom"], 6, replace=False):
        for d in range(3):
            es = []
            for sgn in (1, -1):
                p = s["pos"].copy()
                p[a, d] += sgn * h
                o.set_positions(p)
                o.set_list(idx, lst)
                es.append(ene(o.calforce(abi.MTS_LONG), "pot_tot"))
            fd = -(es[0] - es[1]) / (2 * h)
            assert fd == pytest.approx(f[a, d], rel=2e-5, abs=2e-4), (a, d)
    o.set_positions(s["pos"])


@pytest.mark.parametrize("name", ["water648", "slab_rb"])
def test_virial_is_strain_derivative(oracle_mod, name):
    """W_aa = -dE/d(eps) for x_a -> (1+eps) x_a, L_a -> (1+eps) L_a; per class."""
    s = dict(_systems_with_terms())[name]
    o = oracle_mod.Oracle(s)
    idx, lst = o.mklist(s["rcut_book"], 0)
    r0 = o.calforce(abi.MTS_LONG)
    groups = {
        "pot_virit_lj": ["pot_vdw"],
        "pot_virit_coul": ["pot_elc", "pot_ewk", "pot_ewc", "pot_ewnq"],
        "atm_virit_bond": ["pot_bond"],
        "atm_virit_angl": ["pot_angl"],
        "atm_virit_tors": ["pot_tors", "pot_torsrb"],
        "atm_virit_14": ["pot_vdw14", "pot_elc14"],
    }
    # h = 1e-6: larger steps let pairs cross the cutoff (an energy jump),
    # smaller ones amplify the roundoff of the reciprocal sum (~1e-8 absolute
    # in E, i.e. ~2e-5 of W here); 1e-4 still fails any wrong sign, factor
    # or dropped term
    h = 1e-6
    for ax in range(3):
        es = []
        for sgn in (1, -1):
            sc = np.ones(3)
            sc[ax] += sgn * h
            o.set_box(s["box"] * sc)
            o.set_positions(s["pos"] * sc)
            o.set_list(idx, lst)
            es.append(o.calforce(abi.MTS_LONG))
        for tn, enames in groups.items():
            w = r0["virt"][abi.TENS_NAMES.index(tn)][ax, ax]
            dE = sum(ene(es[0], k) - ene(es[1], k) for k in enames) / (2 * h)
            assert -dE == pytest.approx(w, rel=1e-4, abs=1e-5), (tn, ax)
    o.set_box(s["box"])
    o.set_positions(s["pos"])


# ---------------------------------------------------------------------------
def _hf_system():
    s = systems.propanol_water_slab(nwater=60, nprop=8, box=(16.0, 16.0, 40.0),
                                    nfft=(16, 16, 40), rcut=7.5, rcut_book=8.0, alpha=0.4,
                                    seed=21)
    rng = np.random.default_rng(21)
    s["pos"] = s["pos"] + rng.normal(scale=0.05, size=s["pos"].shape)
    return s


def test_heat_flux_sum_rules(oracle_mod):
    s = _hf_system()
    o = oracle_mod.Oracle(s)
    o.mklist(s["rcut_book"], 0)
    base = o.calforce(abi.MTS_LONG)
    z = np.linspace(0.0, 40.0, 6)
    hf = dict(z1=z[:-1], z2=z[1:])
    cls = {k: i for i, k in enumerate(abi.HF_NAMES)}
    tens = {k: i for i, k in enumerate(abi.TENS_NAMES)}
    bulk = o.calforce_hf(dict(hf, mode=abi.HF_BULK))
    vol = o.calforce_hf(dict(hf, mode=abi.HF_VOLUME))
    pe = bulk["pot_atm"]
    # per-atom energies sum to the class energies
    assert pe[cls["bo"]].sum() == pytest.approx(ene(base, "pot_bond"), rel=1e-10, abs=1e-10)
    assert pe[cls["an"]].sum() == pytest.approx(ene(base, "pot_angl"), rel=1e-10, abs=1e-10)
    assert pe[cls["to"]].sum() == pytest.approx(ene(base, "pot_torsrb") + ene(base, "pot_tors"),
                                                rel=1e-10, abs=1e-10)
    assert pe[cls["14"]].sum() == pytest.approx(ene(base, "pot_vdw14") + ene(base, "pot_elc14"),
                                                rel=1e-10, abs=1e-10)
    assert (pe[cls["ljin"]] + pe[cls["vdw"]]).sum() == pytest.approx(ene(base, "pot_vdw"),
                                                                    rel=1e-10)
    # bulk: per-atom 3x3 virials sum to the class tensors (region 0 holds them)
    w = bulk["viri_atm"]
    wlj = (w[cls["ljin"], 0] + w[cls["vdw"], 0]).sum(0)
    np.testing.assert_allclose(wlj.T, base["virt"][tens["pot_virit_lj"]], rtol=1e-9, atol=1e-9)
    np.testing.assert_allclose(w[cls["bo"], 0].sum(0).T, base["virt"][tens["atm_virit_bond"]],
                               rtol=1e-9, atol=1e-9)
    # volume mode: regions tile [0, Lz], so the z column summed over regions
    # and atoms equals the class tensor's z column
    wv = vol["viri_atm"]
    zlj = (wv[cls["ljin"]] + wv[cls["vdw"]]).sum((0, 1))
    np.testing.assert_allclose(zlj[2, :], base["virt"][tens["pot_virit_lj"]][:, 2],
                               rtol=1e-9, atol=1e-9)
    assert np.all(zlj[:2, :] == 0.0)
    # energies of the volume mode are the same per-atom energies
    np.testing.assert_allclose(vol["pot_atm"], pe, rtol=1e-12, atol=1e-12)
