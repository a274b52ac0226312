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
