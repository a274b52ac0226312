"""pgk_ewaldk_hf on the device against the oracle's restatement of prepewk +
ewaldkp_hf (itself pinned to the compiled reference): the same sums in the
same order, so only sin / cos and contraction differ."""
import numpy as np
import pytest

from helpers import water648

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("netq", [0, 1])
def test_ewaldk_hf_matches_oracle(oracle_mod, netq):
    from thrid_amd.pgk import ForceContext
    s = dict(water648())
    s["charge"] = np.array(s["charge"], np.float64)
    s["charge"][0] += 0.05 * netq
    ref = oracle_mod.Oracle(s).ewaldk_hf(6, ifnetq=netq)
    c = ForceContext(s, "fp64")
    got = c.ewaldk_hf(6, ifnetq=netq)
    c.close()
    for k in ("force", "pot_atm", "viri_atm", "virt"):
        scale = np.max(np.abs(ref[k]))
        assert np.max(np.abs(got[k] - ref[k])) <= 1e-11 * scale, k
    for k in ("pot_ewk", "pot_viri_coul"):
        assert abs(got[k] - ref[k]) <= 1e-12 * abs(ref[k]), k


def test_ewaldk_hf_rejects_bad_input():
    from thrid_amd.pgk import ForceContext, PgkError
    s = water648()
    c = ForceContext(s, "fp64")
    with pytest.raises(PgkError):
        c.ewaldk_hf(0)
    with pytest.raises(PgkError):
        c.ewaldk_hf(4, yratio=0.2)
    c.close()


@pytest.mark.parametrize("prec", ["fp64", "mixed"])
def test_spme_atom_energies(oracle_mod, prec):
    """Per-atom SPME reciprocal energies: they add up to pot_ewk of the same
    call, and with a fine grid they approach the standard Ewald per-atom
    energies (minus the self term) computed by the independent sum."""
    from thrid_amd import abi
    from thrid_amd.pgk import ForceContext, PgkError
    s = dict(water648())
    s["nfft"] = np.array([48, 48, 48], np.int32)
    s["pme_order"] = 8
    c = ForceContext(s, prec)
    with pytest.raises(PgkError):
        c.recip_energy_atom()                 # nothing evaluated yet
    c.mklist(s["rcut_book"])
    r = c.calforce(abi.MTS_LONG, abi.Mts(0, 0, 0, 0, 0, abi.MTS_LONG, 0))
    e = c.recip_energy_atom()
    ewk = r["ene"][abi.ENE_NAMES.index("pot_ewk")]
    tol = 1e-12 if prec == "fp64" else 1e-5
    assert abs(e.sum() - ewk) <= tol * abs(ewk)
    ew = oracle_mod.Oracle(s).ewaldk_hf(10)
    q = np.asarray(s["charge"])
    ref = ew["pot_atm"] + s["alpha"] / np.sqrt(np.pi) * q * q
    assert np.max(np.abs(e - ref)) <= 1e-4 * np.max(np.abs(ref))
    c.set_positions(s["pos"])
    with pytest.raises(PgkError):
        c.recip_energy_atom()                 # the grid belongs to the old call
    c.close()


@pytest.mark.parametrize("prec", ["fp64", "mixed"])
def test_spme_atom_virials(oracle_mod, prec):
    """Per-atom SPME reciprocal virial tensors: symmetric, they add up to the
    reciprocal virial tensor of the same call (bonds and exclusions removed,
    real-space terms off, so the Coulomb tensor is the reciprocal one alone),
    and with a fine grid they approach the standard Ewald per-atom tensors of
    the independent k-space sum (ewaldkp_hf's virielct_atm)."""
    from thrid_amd import abi
    from thrid_amd.pgk import ForceContext, PgkError
    s = dict(water648())
    for k in ("bond_ij", "bond_type", "angl_ijk", "angl_type"):
        s[k] = np.asarray(s[k])[:0]
    s["nfft"] = np.array([48, 48, 48], np.int32)
    s["pme_order"] = 8
    c = ForceContext(s, prec)
    with pytest.raises(PgkError):
        c.recip_virial_atom()                 # nothing evaluated yet
    c.mklist(s["rcut_book"])
    r = c.calforce(abi.MTS_LONG, abi.Mts(0, 0, 0, 0, 0, abi.MTS_LONG, 0))
    w = c.recip_virial_atom()
    vt = r["virt"][abi.TENS_NAMES.index("pot_virit_coul")]
    assert np.max(np.abs(w - np.transpose(w, (0, 2, 1)))) == 0.0
    tol = 1e-10 if prec == "fp64" else 1e-5
    wsum = w.sum(axis=0)
    assert np.max(np.abs(wsum - vt)) <= tol * np.max(np.abs(vt)), (wsum, vt)
    ew = oracle_mod.Oracle(s).ewaldk_hf(10)
    ref = np.asarray(ew["viri_atm"]).reshape(-1, 3, 3)
    err = np.max(np.abs(w - ref)) / np.max(np.abs(ref))
    print(prec, "per-atom virial vs standard Ewald", err)
    assert err <= (1e-7 if prec == "fp64" else 1e-5)
    # the energies of the same evaluation still come from the potential grid
    e = c.recip_energy_atom()
    assert abs(e.sum() - r["ene"][abi.ENE_NAMES.index("pot_ewk")]) <= 1e-5 * abs(e.sum())
    c.set_positions(s["pos"])
    with pytest.raises(PgkError):
        c.recip_virial_atom()
    c.close()
