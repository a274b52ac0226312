"""Exclusion lists: the oracle's mkexcl, the package's generator and the
reference's own mkexcl (mkexcl.F90:6-223, compiled from source) agree bit for
bit, and the oracle matches the reference vectors stored in
tests/golden/mkexcl.npz.

The reference's layout (mkexcl.F90:163-222): 1-based, per atom i the sorted
unique partners j > i taken from the end atoms of bonds, angles, Urey-Bradley
angles, periodic, RB and improper torsions; an atom with no such partner holds a
single 0 entry; index(natom+1) = length + 1.
"""
import os

import numpy as np
import pytest

from thrid_amd import systems
from helpers import water648, small_slab, butane_box

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def _mix_replica():
    return systems.mixture_box(1600, "mix", seed=77, alcohol_frac=0.15, name="mix-excl")


CASES = {
    "water648": water648,
    "slab_rb": lambda: small_slab("rb"),
    "slab_periodic": lambda: small_slab("periodic"),
    "butane_box": butane_box,
    "mix": _mix_replica,
}


def _check_layout(idx, lst, n):
    """Invariants of the reference's final layout, independent of any builder."""
    assert idx[0] == 1 and idx[n] == len(lst) + 1
    for a in range(1, n + 1):
        seg = lst[idx[a - 1] - 1: idx[a] - 1]
        assert len(seg) >= 1
        if len(seg) == 1 and seg[0] == 0:
            continue
        assert np.all(seg > a) and np.all(np.diff(seg) > 0)


@pytest.mark.parametrize("name", list(CASES))
def test_mkexcl_three_ways(oracle_mod, name):
    s = CASES[name]()
    n = int(s["natom"])
    oi, ol = oracle_mod.mkexcl(s)
    gi, gl = systems.exclusions(s)
    _check_layout(oi, ol, n)
    assert np.array_equal(oi, gi) and np.array_equal(ol, gl)
    if not oracle_mod.reference_available():
        pytest.skip("reference not built (oracle/build_ref.sh)")
    ref = oracle_mod.Reference(s, use_mkexcl=True)
    ri, rl = ref.excl()
    assert np.array_equal(ri, oi), "index_excl differs from the reference's mkexcl"
    assert np.array_equal(rl, ol), "list_excl differs from the reference's mkexcl"


def test_mkexcl_golden_vectors(oracle_mod):
    """The oracle on the topologies stored with the reference's own CSR."""
    z = np.load(os.path.join(GOLD, "mkexcl.npz"))
    names = [str(x) for x in z["names"]]
    assert set(names) == set(CASES)
    for nm in names:
        s = {"natom": int(z[f"{nm}_natom"])}
        for k in ("bond_ij", "angl_ijk", "tors_ijkl", "torsrb_ijkl"):
            s[k] = z[f"{nm}_{k}"]
        oi, ol = oracle_mod.mkexcl_topology(s)
        assert np.array_equal(oi, z[f"{nm}_index"]), nm
        assert np.array_equal(ol, z[f"{nm}_list"]), nm
        gi, gl = systems.exclusions(s)
        assert np.array_equal(gi, z[f"{nm}_index"]) and np.array_equal(gl, z[f"{nm}_list"]), nm


def test_mkexcl_pins_catch_planted_mistakes(oracle_mod):
    """Each plausible mistake in building the CSR fails the layout or the
    golden comparison: keeping j < i, dropping the 0 sentinel, keeping
    duplicates, or leaving out the angle (1-3) partners."""
    z = np.load(os.path.join(GOLD, "mkexcl.npz"))
    nm = "slab_rb"
    n = int(z[f"{nm}_natom"])
    ri, rl = z[f"{nm}_index"], z[f"{nm}_list"]
    s = {"natom": n}
    for k in ("bond_ij", "angl_ijk", "tors_ijkl", "torsrb_ijkl"):
        s[k] = z[f"{nm}_{k}"]

    def build(both_dirs=False, sentinel=True, dedup=True, angles=True):
        part = [[] for _ in range(n + 1)]
        pairs = [tuple(p) for p in s["bond_ij"]]
        if angles:
            pairs += [(a, c) for a, _b, c in s["angl_ijk"]]
        pairs += [(a, d) for a, _b, _c, d in s["tors_ijkl"]]
        pairs += [(a, d) for a, _b, _c, d in s["torsrb_ijkl"]]
        for i, j in pairs:
            part[i].append(j)
            part[j].append(i)
        idx, lst = np.zeros(n + 1, np.int32), []
        for a in range(1, n + 1):
            idx[a - 1] = len(lst) + 1
            js = sorted(x for x in part[a] if both_dirs or x > a)
            if dedup:
                js = sorted(set(js))
            lst += js if (js or not sentinel) else [0]
        idx[n] = len(lst) + 1
        return idx, np.array(lst, np.int32)

    good = build()
    assert np.array_equal(good[0], ri) and np.array_equal(good[1], rl)
    for kw in (dict(both_dirs=True), dict(sentinel=False), dict(dedup=False),
               dict(angles=False)):
        bi, bl = build(**kw)
        assert not (np.array_equal(bi, ri) and np.array_equal(bl, rl)), kw
