"""The standard Ewald reciprocal sum with per-atom terms (prepewk +
ewaldkp_hf) in the oracle: pinned to the compiled reference bit for bit, to
its own sum rules (per-atom energies and virials add up to the box values)
and to the independent SPME path, which converges to the same sum."""
import numpy as np
import pytest

from thrid_amd import abi
from helpers import water648


def _charged(s):
    t = dict(s)
    t["charge"] = np.array(s["charge"], np.float64)
    t["charge"][0] += 0.05        # a net charge: the ifnetqcorrp branch matters
    return t


@pytest.mark.parametrize("netq", [0, 1])
def test_oracle_ewaldk_hf_matches_reference(oracle_mod, netq):
    if not oracle_mod.reference_available():
        pytest.skip("reference library not built (oracle/build_ref.sh)")
    s = _charged(water648())
    a = oracle_mod.Reference(s).ewaldk_hf(6, ifnetq=netq)
    b = oracle_mod.Oracle(s).ewaldk_hf(6, ifnetq=netq)
    for k in a:
        np.testing.assert_array_equal(np.asarray(a[k]), np.asarray(b[k]), err_msg=k)
    assert abs(a["pot_ewk"]) > 0 and np.max(np.abs(a["viri_atm"])) > 0


def test_ewaldk_hf_sum_rules(oracle_mod):
    s = water648()
    r = oracle_mod.Oracle(s).ewaldk_hf(7)
    q = np.asarray(s["charge"])
    self_e = -s["alpha"] / np.sqrt(np.pi) * q * q
    np.testing.assert_allclose((r["pot_atm"] - self_e).sum(), r["pot_ewk"], rtol=1e-12)
    np.testing.assert_allclose(r["viri_atm"].sum(0), r["virt"].T, rtol=1e-11,
                               atol=1e-11 * np.max(np.abs(r["virt"])))
    # a neutral box: no net force, and the scalar virial equals the energy
    assert np.max(np.abs(r["force"].sum(0))) < 1e-10 * np.max(np.abs(r["force"]))
    assert r["pot_viri_coul"] == r["pot_ewk"]


def test_ewaldk_hf_agrees_with_spme(oracle_mod):
    """Fine SPME (48^3 grid, order 8) and the Ewald sum to kmax = 10 are two
    routes to the same reciprocal energy, forces and virial tensor."""
    s = dict(water648())
    s["nfft"] = np.array([48, 48, 48], np.int32)
    s["pme_order"] = 8
    o = oracle_mod.Oracle(s)
    o.mklist(s["rcut_book"], 0)
    mts = abi.Mts(0, 0, 0, 0, 0, abi.MTS_LONG, 0)      # the reciprocal term only
    spme = o.calforce(abi.MTS_LONG, mts)
    ew = o.ewaldk_hf(10)
    e_spme = spme["ene"][abi.ENE_NAMES.index("pot_ewk")]
    assert abs(ew["pot_ewk"] - e_spme) < 1e-6 * abs(e_spme)
    f = spme["force"]
    assert np.sqrt(((ew["force"] - f) ** 2).sum() / (f ** 2).sum()) < 1e-5
