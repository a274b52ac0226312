"""The Fortran binding (thrid_amd/fortran/pgk_force_mod.F90) and its example
host program, built by __graft_entry__.build() with flang.

CPU: the bind(C) types have the sizes of the C header's structs and a call
without a device stops with the library's error text.  GPU: the Fortran host
computes the 648-atom water forces through the C ABI and they match the oracle.
"""
import ctypes as C
import os
import subprocess

import numpy as np
import pytest

from thrid_amd import abi
from helpers import rms_rel, water648

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
EXE = os.path.join(ROOT, "thrid_amd", "_build", "fortran", "pgk_example_calforce")

needs_exe = pytest.mark.skipif(not os.path.exists(EXE),
                               reason="Fortran example not built (no flang)")


def write_input(path, s, precision=0):
    from thrid_amd.systems import exclusions
    idx, lst = exclusions(s)
    n, nt = int(s["natom"]), int(s["lj_eps"].shape[0])
    hdr = np.array([n, nt, len(lst), *[int(x) for x in s["nfft"]], int(s["pme_order"]),
                    int(s["nspltbl"]), precision, 0], np.int32)
    par = np.array([s["alpha"], s["rrcut"], s["rcut"], s["rcut_book"], *s["box"]], np.float64)
    with open(path, "wb") as f:
        f.write(hdr.tobytes())
        f.write(par.tobytes())
        f.write(np.asarray(s["charge"], np.float64).tobytes())
        f.write(np.asarray(s["atype"], np.int32).tobytes())
        f.write(np.asarray(s["molid"], np.int32).tobytes())
        f.write(np.asfortranarray(s["lj_eps"]).ravel(order="F").astype(np.float64).tobytes())
        f.write(np.asfortranarray(s["lj_rmin"]).ravel(order="F").astype(np.float64).tobytes())
        f.write(np.asfortranarray(s["inttyp"]).ravel(order="F").astype(np.int32).tobytes())
        f.write(idx.astype(np.int32).tobytes())
        f.write(lst.astype(np.int32).tobytes())
        f.write(np.ascontiguousarray(s["pos"], np.float64).tobytes())


@needs_exe
def test_fortran_types_match_c_layout():
    out = subprocess.run([EXE, "--layout"], check=True, capture_output=True, text=True).stdout
    got = dict((ln.split()[0], int(ln.split()[1])) for ln in out.splitlines() if ln.strip()
               and not ln.startswith("Fortran"))
    assert got["pgk_system_desc"] == C.sizeof(abi.SystemDesc)
    assert got["pgk_mts"] == C.sizeof(abi.Mts)
    assert got["pgk_force_out"] == C.sizeof(abi.ForceOut)
    assert got["pgk_hf_desc"] == C.sizeof(abi.HfDesc)
    assert got["pgk_hf_out"] == C.sizeof(abi.HfOut)
    assert got["abi"] == abi.ABI_VERSION


@needs_exe
def test_fortran_host_reports_missing_device(tmp_path):
    import torch
    if torch.cuda.is_available():
        pytest.skip("a GPU is present")
    inp = tmp_path / "in.bin"
    write_input(str(inp), water648())
    r = subprocess.run([EXE, str(inp), str(tmp_path / "out.bin")], capture_output=True, text=True)
    assert r.returncode != 0
    assert "pgk_create failed" in r.stdout and "device" in r.stdout


@needs_exe
@pytest.mark.gpu
def test_fortran_host_calforce_matches_oracle(tmp_path, oracle_mod):
    s = water648()
    inp, outp = tmp_path / "in.bin", tmp_path / "out.bin"
    write_input(str(inp), s)
    r = subprocess.run([EXE, str(inp), str(outp)], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    raw = np.fromfile(str(outp), np.float64)
    n = int(s["natom"])
    force = raw[:3 * n].reshape(n, 3)
    ene = raw[3 * n:3 * n + abi.NENE]
    o = oracle_mod.Oracle(s)
    o.mklist(s["rcut_book"], 0)
    ref = o.calforce(abi.MTS_LONG)
    assert rms_rel(force, ref["force"]) < 1e-9
    np.testing.assert_allclose(ene, ref["ene"], rtol=1e-9, atol=1e-9)
