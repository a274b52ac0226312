"""pgk_recip_energy_atom_noalloc (fixed device scratch, no allocation) against
pgk_recip_energy_atom, through the C ABI: two calls in a row on one device,
the second on a smaller system, each bit-identical to the allocating entry
point on the same reciprocal evaluation; and the size limit is an error."""
import numpy as np
import pytest

from thrid_amd import abi, systems

pytestmark = pytest.mark.gpu


def _ctx(s, prec):
    from thrid_amd.pgk import ForceContext
    c = ForceContext(s, prec)
    c.mklist(s["rcut_book"])
    c.calforce(abi.MTS_LONG)   # leaves the potential grid of this evaluation
    return c


@pytest.mark.parametrize("prec", ["fp64", "mixed"])
def test_noalloc_matches_allocating_entry_point(prec):
    big = systems.water_box(nmol=1000, box=31.07, nfft=(32, 32, 32), seed=5, name="w3000")
    small = systems.water_box()   # 648 atoms
    assert big["natom"] > small["natom"]
    for s in (big, small):
        c = _ctx(s, prec)
        want = c.recip_energy_atom()
        got = c.recip_energy_atom_noalloc()
        assert got.shape == (s["natom"],)
        assert np.array_equal(got, want), s["name"]
        assert abs(got.sum() - c.calforce(abi.MTS_LONG)["ene"][abi.ENE_NAMES.index("pot_ewk")]) \
            <= 1e-5 * abs(got.sum())
        c.close()


def test_noalloc_second_call_sees_no_stale_entries():
    """A smaller system after a larger one: only its own natom entries are
    returned, and a repeated call gives the same values."""
    big = systems.water_box(nmol=1000, box=31.07, nfft=(32, 32, 32), seed=6, name="w3000b")
    small = systems.water_box(seed=9)
    cb = _ctx(big, "fp64")
    eb = cb.recip_energy_atom_noalloc()
    cs = _ctx(small, "fp64")
    es1 = cs.recip_energy_atom_noalloc()
    es2 = cs.recip_energy_atom_noalloc()
    assert np.array_equal(es1, es2)
    assert np.array_equal(es1, cs.recip_energy_atom())
    assert np.array_equal(eb, cb.recip_energy_atom())
    cb.close()
    cs.close()
