"""The compiled drop-in: the reference itself, built with
thrid_amd/fortran/calforce_gpu.F90 in place of calforce.o, calforce_hf.o,
mklist.o and mklist_cell.o (oracle/build_ref.sh -> oracle/_ref/libpgkref_gpu.so),
so its own harness calls calforce / calforce_hf / mklist* with the reference's
argument lists (interface_mdtech.F90:154-181, 358-951) and reaches the MI355X
library.

CPU: the build links (every caller of the four routines resolves to the shim,
which resolves to libpgkforce.so) and, without a device, a call stops with the
library's error text.  GPU: the shimmed reference matches the plain reference
(compiled from the same sources) on water648 and the water/propanol slab: FP64
forces to 1e-9 rms, energies and virials to 1e-9, calforce_hf's per-atom
arrays (6 volume slabs) to 1e-10 of the largest entry; mixed forces to 1e-4.
"""
import os
import subprocess
import sys

import numpy as np
import pytest

from helpers import rms_rel
from shim_cases import CASES, HF_REGIONS

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
GPU_SO = os.path.join(ROOT, "oracle", "_ref", "libpgkref_gpu.so")

needs_so = pytest.mark.skipif(not os.path.exists(GPU_SO),
                              reason="reference with the shim not built (oracle/build_ref.sh)")


def _run(out, precision, timeout=600, lib=GPU_SO):
    env = dict(os.environ, PGK_REF_SO=lib, PGK_PRECISION=precision)
    return subprocess.run([sys.executable, os.path.join(HERE, "shim_run.py"), str(out)],
                          env=env, capture_output=True, text=True, timeout=timeout)


def _symbols(kind):
    out = subprocess.run(["nm", "-D", kind, GPU_SO], capture_output=True, text=True,
                         check=True).stdout
    return {ln.split()[-1] for ln in out.splitlines() if ln.strip()}


@needs_so
def test_shim_replaces_the_four_routines():
    defined = _symbols("--defined-only")
    for f in ("calforce_", "calforce_hf_", "mklist2a_", "mklist_cell_", "pgkref_calforce"):
        assert f in defined, f
    undefined = _symbols("--undefined-only")
    for f in ("pgk_create", "pgk_calforce", "pgk_calforce_hf", "pgk_build_list"):
        assert f in undefined, f
    # the reference's own force and list routines are not in this build
    plain = os.path.join(ROOT, "oracle", "_ref", "libpgkref.so")
    nm = subprocess.run(["nm", plain], capture_output=True, text=True, check=True).stdout
    assert "ewaldrp_" in nm
    nm_gpu = subprocess.run(["nm", GPU_SO], capture_output=True, text=True, check=True).stdout
    assert "pgk_calforce" in nm_gpu


@needs_so
def test_shim_without_device_stops_with_library_error(tmp_path):
    import torch
    if torch.cuda.is_available():
        pytest.skip("a device is present")
    r = _run(tmp_path / "x.npz", "fp64", timeout=300)
    assert "SHIM_RUN_DONE" not in r.stdout
    assert "calforce_gpu: pgk_create" in r.stdout + r.stderr


@pytest.mark.gpu
@needs_so
@pytest.mark.timeout(600)
@pytest.mark.parametrize("precision,ftol", [("fp64", 1e-9), ("mixed", 1e-4)])
def test_shim_matches_reference(tmp_path, precision, ftol):
    # both builds run in fresh interpreters of their own (the harness library
    # is a process-wide singleton)
    plain = os.path.join(ROOT, "oracle", "_ref", "libpgkref.so")
    rr = _run(tmp_path / "plain.npz", "fp64", lib=plain)
    assert rr.returncode == 0 and "SHIM_RUN_DONE" in rr.stdout, rr.stdout[-2000:]
    r = _run(tmp_path / f"shim_{precision}.npz", precision)
    print(r.stdout[-2000:], r.stderr[-2000:])
    assert r.returncode == 0 and "SHIM_RUN_DONE" in r.stdout
    want = np.load(tmp_path / "plain.npz")
    got = np.load(tmp_path / f"shim_{precision}.npz")
    for name in CASES:
        err = rms_rel(got[f"{name}_force"], want[f"{name}_force"])
        print(f"{name} {precision}: rms rel force err {err:.3e}")
        assert err <= ftol, (name, err)
        etol = 1e-9 if precision == "fp64" else 1e-4
        np.testing.assert_allclose(got[f"{name}_ene"], want[f"{name}_ene"], rtol=etol, atol=etol)
        vs = max(1.0, float(np.max(np.abs(want[f"{name}_virt"]))))
        assert float(np.max(np.abs(got[f"{name}_virt"] - want[f"{name}_virt"]))) <= \
            (1e-9 if precision == "fp64" else 1e-3) * vs
        if name in HF_REGIONS:
            assert rms_rel(got[f"{name}_hf_force"], want[f"{name}_hf_force"]) <= ftol
            pe, w = want[f"{name}_hf_pot_atm"], want[f"{name}_hf_viri_atm"]
            tol = 1e-10 if precision == "fp64" else 1e-4
            assert np.max(np.abs(got[f"{name}_hf_pot_atm"] - pe)) <= \
                tol * max(1.0, np.max(np.abs(pe)))
            assert np.max(np.abs(got[f"{name}_hf_viri_atm"] - w)) <= \
                tol * max(1.0, np.max(np.abs(w)))
