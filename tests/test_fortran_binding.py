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


FLANG = "/opt/rocm/lib/llvm/bin/flang"
MODDIR = os.path.join(ROOT, "thrid_amd", "_build", "fortran")


@pytest.mark.skipif(not (os.path.exists(FLANG) and
                         os.path.exists(os.path.join(MODDIR, "pgk_force_mod.mod"))),
                    reason="flang or the built module missing")
def test_fortran_host_writes_out_htf_lines(tmp_path):
    """A Fortran host writes out_htf.dat lines through the module's
    interfaces (host-only calls, no device): the same text as the C ABI
    gives Python, with hf in the reference's (3, PGK_NHFJ, nregion) order."""
    from thrid_amd import formats
    src = tmp_path / "htf.F90"
    src.write_text("""
program htf
  use iso_c_binding
  use pgk_force_mod
  implicit none
  integer, parameter :: nreg = 2
  real(c_double) :: hf(3, PGK_NHFJ, nreg)
  character(kind=c_char) :: buf(20000)
  integer :: i, n, t, r
  do r = 1, nreg
    do t = 1, PGK_NHFJ
      do n = 1, 3
        hf(n, t, r) = r * 1000.0d0 + t * 10.0d0 + n + 0.125d0
      end do
    end do
  end do
  if (pgk_out_htf_header(nreg, 1, buf, 20000) /= 0) stop 1
  call put(buf)
  if (pgk_out_htf_row(5, hf, nreg, 1, 0, 0.5d0, buf, 20000) /= 0) stop 2
  call put(buf)
contains
  subroutine put(b)
    character(kind=c_char), intent(in) :: b(*)
    integer :: k
    k = 1
    do while (b(k) /= c_null_char)
      k = k + 1
    end do
    write(*, '(*(A))', advance='no') (b(i), i = 1, k - 1)
  end subroutine put
end program htf
""")
    exe = tmp_path / "htf"
    # the program calls only bind(C) interfaces of the module: no module object to link
    subprocess.run([FLANG, "-O1", "-I", MODDIR, "-o", str(exe), str(src),
                    "-L" + os.path.join(ROOT, "thrid_amd"), "-lpgkforce",
                    "-Wl,-rpath," + os.path.join(ROOT, "thrid_amd")],
                   check=True, capture_output=True)
    out = subprocess.run([str(exe)], check=True, capture_output=True, text=True).stdout
    hf = np.zeros((2, abi.NHFJ, 3))
    for r in range(2):
        for t in range(abi.NHFJ):
            for n in range(3):
                hf[r, t, n] = (r + 1) * 1000.0 + (t + 1) * 10.0 + (n + 1) + 0.125
    want = formats.out_htf_header(2, True) + formats.out_htf_row(5, hf, jref=0.5, all_dir=True)
    assert out == want
