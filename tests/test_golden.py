"""Oracle against golden vectors written by the reference (tests/golden/*.npz,
produced by tests/golden/make_golden.py from peachgk_md compiled from source).

The reference ships no tests or force fixtures of its own; these vectors are
its outputs on seeded synthetic inputs that are stored inside the fixture.
"""
import os

import numpy as np
import pytest

from thrid_amd import abi
from helpers import rms_rel

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def load(name):
    z = np.load(os.path.join(GOLD, name + ".npz"))
    s = {}
    for k in z.files:
        if k.startswith("sys_"):
            v = z[k]
            s[k[4:]] = v.item() if v.ndim == 0 else v
    s["natom"] = int(s["natom"])
    s["nmol"] = int(s["nmol"])
    s["nspltbl"] = int(s["nspltbl"])
    s["pme_order"] = int(s["pme_order"])
    return s, z


@pytest.mark.parametrize("name", ["water648", "slab_rb", "slab_periodic"])
def test_oracle_matches_reference_vectors(oracle_mod, name):
    s, z = load(name)
    o = oracle_mod.Oracle(s)
    idx, lst = o.mklist(s["rcut_book"], int(z["list_mode"]))
    assert np.array_equal(idx, z["nb_index"]) and np.array_equal(lst, z["nb_list"])
    r = o.calforce(abi.MTS_LONG)
    assert rms_rel(r["force"], z["force"]) < 1e-12
    assert rms_rel(r["for_viri_coul"], z["for_viri_coul"]) < 1e-12
    assert rms_rel(r["for_viri_lj"], z["for_viri_lj"]) < 1e-12
    np.testing.assert_allclose(r["ene"], z["ene"], rtol=1e-11, atol=1e-9)
    np.testing.assert_allclose(r["virt"], z["virt"], rtol=1e-11, atol=1e-8)


@pytest.mark.parametrize("mode", [abi.HF_VOLUME, abi.HF_SURFACE])
def test_oracle_heat_flux_vectors(oracle_mod, mode):
    s, z = load("hf_slab")
    o = oracle_mod.Oracle(s)
    o.mklist(s["rcut_book"], 0)
    zz = z["z"]
    r = o.calforce_hf(dict(z1=zz[:-1], z2=zz[1:], mode=mode))
    assert rms_rel(r["force"], z[f"m{mode}_force"]) < 1e-12
    np.testing.assert_allclose(r["ene"], z[f"m{mode}_ene"], rtol=1e-11, atol=1e-9)
    np.testing.assert_allclose(r["pot_atm"], z[f"m{mode}_pot_atm"], rtol=1e-11, atol=1e-10)
    w = z[f"m{mode}_viri_atm"]
    assert np.max(np.abs(r["viri_atm"] - w)) <= 1e-11 * max(1.0, np.max(np.abs(w)))
