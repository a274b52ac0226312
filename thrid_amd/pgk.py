"""Python binding of libpgkforce.so (the MI355X force path).

Mirrors the reference's force interface: a ForceContext owns one system and
offers mklist (mklist2a / mklist_cell), calforce (calforce.F90:8-1073) and
calforce_hf (calforce_hf.F90:8-1451) with the reference's output records.
There is no CPU fallback: if the HIP library is missing or no GPU is present
every call raises.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

from . import abi

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libpgkforce.so")

_lib = None


class PgkError(RuntimeError):
    pass


def load_library(path: str = LIB_PATH):
    """Load libpgkforce.so and declare its C signatures."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(path):
        raise PgkError(f"libpgkforce.so not built ({path}); run __graft_entry__.build()")
    L = C.CDLL(path)
    vp = C.c_void_p
    L.pgk_abi_version.restype = C.c_int32
    L.pgk_last_error.restype = C.c_char_p
    L.pgk_create.argtypes = [C.POINTER(abi.SystemDesc), C.c_int32, C.POINTER(vp)]
    L.pgk_destroy.argtypes = [vp]
    L.pgk_set_box.argtypes = [vp, vp]
    L.pgk_set_positions.argtypes = [vp, vp]
    L.pgk_set_positions_device.argtypes = [vp, vp]
    L.pgk_build_list.argtypes = [vp, C.c_double]
    L.pgk_list_npairs.argtypes = [vp, C.POINTER(C.c_int64)]
    L.pgk_list_pairs.argtypes = [vp, vp, C.c_int64]
    L.pgk_calforce.argtypes = [vp, C.c_int32, C.POINTER(abi.Mts), C.c_int32,
                               C.POINTER(abi.ForceOut)]
    L.pgk_calforce_hf.argtypes = [vp, C.c_int32, C.POINTER(abi.Mts), C.POINTER(abi.HfDesc),
                                  C.POINTER(abi.ForceOut), C.POINTER(abi.HfOut)]
    L.pgk_calforce_device.argtypes = [vp, C.c_int32, C.POINTER(abi.Mts), C.c_int32, vp, vp, vp]
    L.pgk_calforce_async.argtypes = [vp, C.c_int32, C.POINTER(abi.Mts), C.c_int32, vp]
    L.pgk_energies.argtypes = [vp, vp, vp]
    L.pgk_synchronize.argtypes = [vp]
    L.pgk_md_setup.argtypes = [vp, C.POINTER(abi.MdDesc)]
    L.pgk_md_set_velocities.argtypes = [vp, vp]
    L.pgk_md_get_state.argtypes = [vp, vp, vp]
    L.pgk_md_run.argtypes = [vp, C.c_int32, C.c_int32, C.c_double, C.c_int32, vp]
    L.pgk_md_use_graphs.argtypes = [vp, C.c_int32]
    L.pgk_calheatf.argtypes = [vp, vp, vp, vp]
    L.pgk_set_atmvel_strm.argtypes = [vp, vp]
    L.pgk_ewaldk_hf.argtypes = [vp, C.c_int32, C.c_double, C.c_double, C.c_int32] + [vp] * 6
    L.pgk_calforce_heatf.argtypes = [vp, C.c_int32, C.POINTER(abi.Mts), C.POINTER(abi.HfDesc),
                                     vp, vp, C.POINTER(abi.ForceOut), vp]
    L.pgk_md_nhc.argtypes = [vp, C.POINTER(abi.NhcDesc)]
    L.pgk_md_respa.argtypes = [vp, C.c_int32, C.POINTER(abi.Mts)]
    L.pgk_recip_energy_atom.argtypes = [vp, vp]
    L.pgk_recip_energy_atom_noalloc.argtypes = [vp, vp]
    L.pgk_recip_virial_atom.argtypes = [vp, vp]
    L.pgk_md_heatf.argtypes = [vp, C.POINTER(abi.HfDesc), C.c_int32]
    L.pgk_md_heatf_series.argtypes = [vp, vp, C.c_int32, vp]
    L.pgk_md_nhc_state.argtypes = [vp, vp, vp]
    L.pgk_md_langevin.argtypes = [vp, C.POINTER(abi.LangevinDesc)]
    L.pgk_md_langevin_state.argtypes = [vp, vp, vp]
    L.pgk_md_run_batch.argtypes = [vp, C.c_int32, C.c_int32, C.c_int32, vp, C.c_int32, vp]
    L.pgk_profile.argtypes = [vp, C.c_int32, vp, vp]
    L.pgk_list_info.argtypes = [vp, vp]
    L.pgk_profile_stages.argtypes = [vp, C.c_uint32]
    L.pgk_tune.argtypes = [vp, C.c_int32, C.c_int64]
    for f in ("pgk_create", "pgk_destroy", "pgk_set_box", "pgk_set_positions",
              "pgk_set_positions_device", "pgk_build_list", "pgk_list_npairs",
              "pgk_list_pairs", "pgk_calforce", "pgk_calforce_hf", "pgk_calforce_device",
              "pgk_calforce_async", "pgk_energies", "pgk_synchronize", "pgk_profile",
              "pgk_list_info", "pgk_tune", "pgk_md_setup", "pgk_md_set_velocities", "pgk_md_get_state",
              "pgk_md_run", "pgk_md_use_graphs", "pgk_md_run_batch", "pgk_profile_stages",
              "pgk_md_nhc", "pgk_md_nhc_state", "pgk_calheatf", "pgk_calforce_heatf",
              "pgk_set_atmvel_strm",
              "pgk_md_heatf", "pgk_md_heatf_series", "pgk_ewaldk_hf",
              "pgk_md_respa", "pgk_recip_energy_atom", "pgk_recip_energy_atom_noalloc", "pgk_recip_virial_atom", "pgk_md_langevin",
              "pgk_md_langevin_state"):
        getattr(L, f).restype = C.c_int
    if L.pgk_abi_version() != abi.ABI_VERSION:
        raise PgkError("libpgkforce.so ABI version mismatch")
    _lib = L
    return L


def _check(rc, what):
    if rc != 0:
        msg = _lib.pgk_last_error().decode(errors="replace") if _lib else ""
        raise PgkError(f"{what} failed ({rc}): {msg}")


_TUNED = []


def _tuned_chunk_below():
    """The pair-form crossover of the tuning table written by
    thrid_amd.autotune (None without a table)."""
    if not _TUNED:
        from thrid_amd import autotune
        t = autotune.load_table()
        _TUNED.append(None if t is None else int(t["pair_chunk_below"]))
    return _TUNED[0]


class ForceContext:
    """One system on the current GPU.

    precision: 'fp64' (the reference's arithmetic), 'mixed' (FP32 pair/FFT,
    64-bit fixed-point force sums) or 'fp32'."""

    def __init__(self, sysd, precision="fp64", excl=None, ifnetqcorrp=0, tuned=True):
        self.L = load_library()
        self.sysd = sysd
        self.n = int(sysd["natom"])
        self.precision = precision
        self._desc, self._keep = abi.make_desc(sysd, excl, ifnetqcorrp)
        h = C.c_void_p()
        _check(self.L.pgk_create(C.byref(self._desc), abi.PREC[precision], C.byref(h)),
               "pgk_create")
        self.h = h
        if tuned:
            below = _tuned_chunk_below()
            if below is not None and below != abi.TUNE_DEFAULT_CHUNK_BELOW:
                self.tune(abi.TUNE_PAIR_CHUNK_BELOW, below)
        self.set_box(sysd["box"])
        self.set_positions(sysd["pos"])

    def close(self):
        if getattr(self, "h", None):
            self.L.pgk_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_box(self, box):
        b = np.ascontiguousarray(box, np.float64)
        _check(self.L.pgk_set_box(self.h, b.ctypes.data), "pgk_set_box")

    def set_positions(self, pos):
        p = np.ascontiguousarray(pos, np.float64)
        if p.shape != (self.n, 3):
            raise ValueError("positions must be (natom, 3)")
        _check(self.L.pgk_set_positions(self.h, p.ctypes.data), "pgk_set_positions")

    def set_positions_device(self, dev_ptr: int):
        _check(self.L.pgk_set_positions_device(self.h, C.c_void_p(dev_ptr)),
               "pgk_set_positions_device")

    def mklist(self, rcut_book=None):
        rcb = float(rcut_book if rcut_book is not None else self.sysd["rcut_book"])
        _check(self.L.pgk_build_list(self.h, rcb), "pgk_build_list")

    def npairs(self):
        v = C.c_int64()
        _check(self.L.pgk_list_npairs(self.h, C.byref(v)), "pgk_list_npairs")
        return v.value

    def list_pairs(self):
        """Listed pairs as a sorted (npairs, 2) array of 1-based (i, j), i < j."""
        n = self.npairs()
        out = np.zeros((max(n, 1), 2), np.int32)
        _check(self.L.pgk_list_pairs(self.h, out.ctypes.data, n), "pgk_list_pairs")
        return out[:n]

    def calforce(self, mts_flag=abi.MTS_LONG, mts=None, ifcalpreatom=1, split=False, ene=None):
        mts = mts or abi.mts_all(mts_flag)
        f = np.zeros((self.n, 3))
        e = np.zeros(abi.NENE) if ene is None else np.array(ene, np.float64)
        v = np.zeros((abi.NTENS, 9))
        fc = np.zeros((self.n, 3)) if split else None
        fl = np.zeros((self.n, 3)) if split else None
        o = abi.ForceOut(abi.dptr(f), abi.dptr(e), abi.dptr(v), abi.dptr(fc), abi.dptr(fl))
        _check(self.L.pgk_calforce(self.h, mts_flag, C.byref(mts), ifcalpreatom, C.byref(o)),
               "pgk_calforce")
        r = dict(force=f, ene=e, virt=v.reshape(abi.NTENS, 3, 3).transpose(0, 2, 1).copy())
        if split:
            r["for_viri_coul"], r["for_viri_lj"] = fc, fl
        return r

    def calforce_device(self, d_force=0, mts_flag=abi.MTS_LONG, mts=None, ifcalpreatom=1):
        mts = mts or abi.mts_all(mts_flag)
        e = np.zeros(abi.NENE)
        v = np.zeros((abi.NTENS, 9))
        _check(self.L.pgk_calforce_device(self.h, mts_flag, C.byref(mts), ifcalpreatom,
                                          C.c_void_p(d_force), e.ctypes.data, v.ctypes.data),
               "pgk_calforce_device")
        return e, v

    def calforce_async(self, d_force=0, mts_flag=abi.MTS_LONG, mts=None, ifcalpreatom=1):
        """Queue one force call without waiting; read energies with energies()."""
        mts = mts or abi.mts_all(mts_flag)
        self._mts_keep = mts
        _check(self.L.pgk_calforce_async(self.h, mts_flag, C.byref(mts), ifcalpreatom,
                                         C.c_void_p(d_force)), "pgk_calforce_async")

    def energies(self, ene=None):
        """Wait, then the energy record and virial tensors of the last call."""
        e = np.zeros(abi.NENE) if ene is None else np.array(ene, np.float64)
        v = np.zeros((abi.NTENS, 9))
        _check(self.L.pgk_energies(self.h, e.ctypes.data, v.ctypes.data), "pgk_energies")
        return e, v.reshape(abi.NTENS, 3, 3).transpose(0, 2, 1).copy()

    def _hf_desc(self, hf, keep):
        h = abi.HfDesc()
        h.mode = int(hf.get("mode", abi.HF_VOLUME))
        for k in ("z1", "z2"):
            keep[k] = np.ascontiguousarray(hf[k], np.float64)
        h.nregion = len(keep["z1"])
        h.z1, h.z2 = abi.dptr(keep["z1"]), abi.dptr(keep["z2"])
        keep["ht"] = np.ascontiguousarray(hf.get("hftyp", np.ones(self.n, np.int32)), np.int32)
        h.hftyp = abi.iptr(keep["ht"])
        keep["mc"] = np.ascontiguousarray(hf.get("molecom", np.zeros((self.sysd["nmol"], 3))),
                                          np.float64)
        h.nmol = keep["mc"].shape[0]
        h.molecom = abi.dptr(keep["mc"])
        return h

    def calforce_hf(self, hf, mts_flag=abi.MTS_LONG, mts=None, per_atom=True):
        """calforce_hf; per_atom=False leaves the per-atom arrays on the
        device (for calheatf) and returns None in their place."""
        mts = mts or abi.mts_all(mts_flag)
        keep = {}
        h = self._hf_desc(hf, keep)
        f = np.zeros((self.n, 3))
        e = np.zeros(abi.NENE)
        v = np.zeros((abi.NTENS, 9))
        o = abi.ForceOut(abi.dptr(f), abi.dptr(e), abi.dptr(v), None, None)
        if per_atom:
            pe = np.zeros((abi.NHF, self.n))
            wv = np.zeros((abi.NHF, h.nregion, self.n, 9))
            ho = C.byref(abi.HfOut(abi.dptr(pe), abi.dptr(wv)))
        else:
            pe = wv = ho = None
        _check(self.L.pgk_calforce_hf(self.h, mts_flag, C.byref(mts), C.byref(h), C.byref(o),
                                      ho), "pgk_calforce_hf")
        self._hf_nreg = h.nregion
        return dict(force=f, ene=e, virt=v.reshape(abi.NTENS, 3, 3).transpose(0, 2, 1).copy(),
                    pot_atm=pe,
                    viri_atm=None if wv is None else wv.reshape(abi.NHF, h.nregion, self.n, 3, 3))

    def calforce_heatf(self, hf, vel, mass=None, mts_flag=abi.MTS_LONG, mts=None):
        """calforce_hf + calheatf in one pass without the per-atom arrays:
        returns the calforce record (force, ene, virt) and hf[nregion,
        PGK_NHFJ, 3]."""
        mts = mts or abi.mts_all(mts_flag)
        keep = {}
        h = self._hf_desc(hf, keep)
        f = np.zeros((self.n, 3))
        e = np.zeros(abi.NENE)
        v = np.zeros((abi.NTENS, 9))
        o = abi.ForceOut(abi.dptr(f), abi.dptr(e), abi.dptr(v), None, None)
        vv = np.ascontiguousarray(vel, np.float64)
        mm = None if mass is None else np.ascontiguousarray(mass, np.float64)
        out = np.zeros((h.nregion, abi.NHFJ, 3))
        _check(self.L.pgk_calforce_heatf(self.h, mts_flag, C.byref(mts), C.byref(h),
                                         vv.ctypes.data, None if mm is None else mm.ctypes.data,
                                         C.byref(o), out.ctypes.data), "pgk_calforce_heatf")
        return dict(force=f, ene=e, virt=v.reshape(abi.NTENS, 3, 3).transpose(0, 2, 1).copy(),
                    hf=out)

    def ewaldk_hf(self, kmax, yratio=1.0, zratio=1.0, ifnetq=0):
        """Standard Ewald reciprocal sum with per-atom terms (prepewk +
        ewaldkp_hf) from zeroed outputs: force, pot_ewk, pot_viri_coul, virt
        (force comp, r comp), pot_atm, viri_atm (natom, r comp n, force comp a)."""
        n = self.n
        f, pe, wv = np.zeros((n, 3)), np.zeros(n), np.zeros((n, 9))
        ewk, pvc, vt = np.zeros(1), np.zeros(1), np.zeros(9)
        _check(self.L.pgk_ewaldk_hf(self.h, int(kmax), float(yratio), float(zratio), int(ifnetq),
                                    f.ctypes.data, ewk.ctypes.data, pvc.ctypes.data,
                                    vt.ctypes.data, pe.ctypes.data, wv.ctypes.data),
               "pgk_ewaldk_hf")
        return dict(force=f, pot_ewk=ewk[0], pot_viri_coul=pvc[0], virt=vt.reshape(3, 3).T.copy(),
                    pot_atm=pe, viri_atm=wv.reshape(n, 3, 3))

    def calheatf(self, vel, mass=None):
        """calheatf on the device over the last calforce_hf call's arrays:
        returns hf[nregion, PGK_NHFJ, 3] (abi.HFJ_NAMES).  mass=None reuses
        the last masses given."""
        vv = np.ascontiguousarray(vel, np.float64)
        mm = None if mass is None else np.ascontiguousarray(mass, np.float64)
        out = np.zeros((getattr(self, "_hf_nreg", 1), abi.NHFJ, 3))
        _check(self.L.pgk_calheatf(self.h, vv.ctypes.data, None if mm is None else mm.ctypes.data,
                                   out.ctypes.data), "pgk_calheatf")
        return out

    def set_atmvel_strm(self, vstrm):
        """Streaming velocities [natom, 3] (atmvel_strm) that calheatf and
        calforce_heatf subtract for the kinetic energies; None clears them."""
        if vstrm is None:
            _check(self.L.pgk_set_atmvel_strm(self.h, None), "pgk_set_atmvel_strm")
            return
        vs = np.ascontiguousarray(vstrm, np.float64)
        if vs.shape != (self.n, 3):
            raise ValueError("set_atmvel_strm needs [natom, 3]")
        _check(self.L.pgk_set_atmvel_strm(self.h, vs.ctypes.data), "pgk_set_atmvel_strm")

    # ---- device-resident NVE step ------------------------------------------
    def md_setup(self, mass, cons_ij, cons_d, group_start, dt, tol=1e-10):
        """mass [natom] amu; cons_ij (ncons,2) 1-based; cons_d (ncons,) A;
        group_start (ngroup+1,) 0-based offsets; dt in fs."""
        k = self._md_keep = dict(m=np.ascontiguousarray(mass, np.float64),
                                 ij=np.ascontiguousarray(cons_ij, np.int32).reshape(-1),
                                 d=np.ascontiguousarray(cons_d, np.float64),
                                 g=np.ascontiguousarray(group_start, np.int32))
        d = abi.MdDesc(abi.dptr(k["m"]), len(k["d"]), abi.iptr(k["ij"]), abi.dptr(k["d"]),
                       max(len(k["g"]) - 1, 0), abi.iptr(k["g"]), float(dt), float(tol))
        _check(self.L.pgk_md_setup(self.h, C.byref(d)), "pgk_md_setup")

    def md_set_velocities(self, vel):
        v = np.ascontiguousarray(vel, np.float64)
        _check(self.L.pgk_md_set_velocities(self.h, v.ctypes.data), "pgk_md_set_velocities")

    def md_state(self):
        p = np.zeros((self.n, 3))
        v = np.zeros((self.n, 3))
        _check(self.L.pgk_md_get_state(self.h, p.ctypes.data, v.ctypes.data), "pgk_md_get_state")
        return p, v

    def md_run(self, nstep, nstep_book, rcut_book, nsample):
        """Velocity Verlet + RATTLE; returns (nstep//nsample + 1, 3) rows of
        kinetic, potential and total energy (eref)."""
        out = np.zeros((nstep // nsample + 1, 3))
        _check(self.L.pgk_md_run(self.h, int(nstep), int(nstep_book), float(rcut_book),
                                 int(nsample), out.ctypes.data), "pgk_md_run")
        return out

    def md_use_graphs(self, single=True, batch=True):
        """Replay plain MD steps from a captured HIP graph: `single` for md_run,
        `batch` for this replica in md_run_batch (both off by default)."""
        _check(self.L.pgk_md_use_graphs(self.h, int(bool(single)) | (int(bool(batch)) << 1)),
               "pgk_md_use_graphs")

    def md_nhc(self, temperature, tfreq=None, nfree=None, mchain=4, nyosh=3, next=1,
               xlogs=None, vlogs=None):
        """Nose-Hoover chain thermostat for md_run / md_run_batch
        (moldyn_nhc.F90 with every term on the long step): temperature in K,
        tfreq in 1/fs, nfree degrees of freedom; temperature=None switches it
        off.  With the chain on, the total-energy column of md_run is the
        conserved quantity."""
        if temperature is None:
            _check(self.L.pgk_md_nhc(self.h, None), "pgk_md_nhc")
            self._nhc_m = 0
            return
        if tfreq is None or nfree is None:
            raise ValueError("md_nhc needs tfreq and nfree")
        k = self._nhc_keep = dict(
            x=None if xlogs is None else np.ascontiguousarray(xlogs, np.float64),
            v=None if vlogs is None else np.ascontiguousarray(vlogs, np.float64))
        for a in k.values():
            if a is not None and a.shape != (mchain,):
                raise ValueError("xlogs / vlogs need mchain entries")
        d = abi.NhcDesc(int(mchain), float(temperature), float(tfreq), float(nfree), int(nyosh),
                        int(next), abi.dptr(k["x"]), abi.dptr(k["v"]))
        _check(self.L.pgk_md_nhc(self.h, C.byref(d)), "pgk_md_nhc")
        self._nhc_m = int(mchain)

    def recip_energy_atom(self):
        """Per-atom SPME reciprocal energies of the last reciprocal-space
        evaluation at the current positions (sum = pot_ewk)."""
        e = np.zeros(self.n)
        _check(self.L.pgk_recip_energy_atom(self.h, e.ctypes.data), "pgk_recip_energy_atom")
        return e

    def recip_energy_atom_noalloc(self):
        """recip_energy_atom through the entry point that allocates nothing
        (pgk_recip_energy_atom_noalloc: fixed device scratch)."""
        e = np.zeros(self.n)
        _check(self.L.pgk_recip_energy_atom_noalloc(self.h, e.ctypes.data),
               "pgk_recip_energy_atom_noalloc")
        return e

    def recip_virial_atom(self):
        """Per-atom SPME reciprocal virial tensors (natom, 3, 3) of the last
        reciprocal-space evaluation at the current positions (their sum is
        the reciprocal virial tensor)."""
        w = np.zeros((self.n, 3, 3))
        _check(self.L.pgk_recip_virial_atom(self.h, w.ctypes.data), "pgk_recip_virial_atom")
        return w

    def md_respa(self, nstep_short, mts=None):
        """RESPA for md_run / md_run_batch: nstep_short inner steps per
        outer step, mts = abi.Mts with each term's level (None: all long)."""
        _check(self.L.pgk_md_respa(self.h, int(nstep_short),
                                   None if mts is None else C.byref(mts)), "pgk_md_respa")

    def md_heatf(self, hf, interval):
        """Heat-flux rows every `interval` MD steps (and at step 0): hf as
        for calforce_hf (every atom its own centre); hf=None turns it off."""
        if hf is None:
            _check(self.L.pgk_md_heatf(self.h, None, 0), "pgk_md_heatf")
            return
        keep = self._md_hf_keep = {}
        h = self._hf_desc(hf, keep)
        _check(self.L.pgk_md_heatf(self.h, C.byref(h), int(interval)), "pgk_md_heatf")
        self._md_hf_nreg = h.nregion

    def md_heatf_series(self):
        """Rows of the last md_run: (nrows, nregion, PGK_NHFJ, 3)."""
        nreg = getattr(self, "_md_hf_nreg", 1)
        n = C.c_int32(0)
        _check(self.L.pgk_md_heatf_series(self.h, None, 0, C.byref(n)), "pgk_md_heatf_series")
        out = np.zeros((max(n.value, 1), nreg, abi.NHFJ, 3))
        _check(self.L.pgk_md_heatf_series(self.h, out.ctypes.data, n.value, C.byref(n)),
               "pgk_md_heatf_series")
        return out[:n.value]

    def md_langevin(self, regions, temperature=None, damp=None, mol_com=None, seed=0,
                    stream=None):
        """Region Langevin thermostat for md_run / md_run_batch
        (langevin_regist.F90 + langevin_region.F90): regions [nregion, 6] of
        (x1, x2, y1, y2, z1, z2) in A, temperature [nregion] in K, damp
        [nregion] in 1/fs, mol_com [nmol] 1 where a molecule is registered
        by its centre of mass (None: every molecule); stream = (z1, z2,
        vel [nslab, 3] in A/fs) streaming-velocity slabs (ass_strmvel) or
        None; regions=None switches it off, regions=[] with stream sets the
        streaming slabs alone (no thermostat)."""
        if regions is None:
            _check(self.L.pgk_md_langevin(self.h, None), "pgk_md_langevin")
            self._lt_nreg = 0
            return
        reg = np.asarray(regions, np.float64)
        reg = reg.reshape(0, 6) if reg.size == 0 else np.ascontiguousarray(np.atleast_2d(reg))
        nr = reg.shape[0]
        if reg.shape[1] != 6 or (nr > 0 and (temperature is None or damp is None)):
            raise ValueError("md_langevin needs regions [nregion, 6], temperature and damp")
        cols = [np.ascontiguousarray(reg[:, k]) if nr else None for k in range(6)]
        t = g = None
        if nr:
            t = np.ascontiguousarray(np.broadcast_to(np.asarray(temperature, np.float64), (nr,)))
            g = np.ascontiguousarray(np.broadcast_to(np.asarray(damp, np.float64), (nr,)))
        mc = None if mol_com is None else np.ascontiguousarray(mol_com, np.int32)
        sz1 = sz2 = sv = None
        if stream is not None:
            sz1 = np.ascontiguousarray(stream[0], np.float64)
            sz2 = np.ascontiguousarray(stream[1], np.float64)
            sv = np.ascontiguousarray(stream[2], np.float64).reshape(-1)
            if not (len(sz1) == len(sz2) and sv.size == 3 * len(sz1)):
                raise ValueError("stream needs z1, z2 [nslab] and vel [nslab, 3]")
        self._lt_keep = (cols, t, g, mc, sz1, sz2, sv)
        d = abi.LangevinDesc(nr, *[abi.dptr(a) for a in cols], abi.dptr(t), abi.dptr(g),
                             abi.iptr(mc), int(seed) & 0xFFFFFFFFFFFFFFFF,
                             0 if sz1 is None else len(sz1), abi.dptr(sz1), abi.dptr(sz2),
                             abi.dptr(sv))
        _check(self.L.pgk_md_langevin(self.h, C.byref(d)), "pgk_md_langevin")
        self._lt_nreg = nr

    def md_langevin_state(self):
        """(det_ene_kin [nregion] in eref over the last run, natom_lt
        [nregion] at the last registration)."""
        nr = getattr(self, "_lt_nreg", 0)
        det = np.zeros(max(nr, 1))
        cnt = np.zeros(max(nr, 1), np.int32)
        _check(self.L.pgk_md_langevin_state(self.h, det.ctypes.data, cnt.ctypes.data),
               "pgk_md_langevin_state")
        return det[:nr], cnt[:nr]

    def md_nhc_state(self):
        """(xlogs, vlogs) of the chain, vlogs in 1/fs."""
        m = getattr(self, "_nhc_m", 0)
        x = np.zeros(max(m, 1))
        v = np.zeros(max(m, 1))
        _check(self.L.pgk_md_nhc_state(self.h, x.ctypes.data, v.ctypes.data), "pgk_md_nhc_state")
        return x[:m], v[:m]

    def list_info(self):
        """(clusters, list entries, listed atom pairs, pair slots per entry)."""
        v = np.zeros(4, np.int64)
        _check(self.L.pgk_list_info(self.h, v.ctypes.data), "pgk_list_info")
        return tuple(int(x) for x in v)

    def profile(self, enable=-1):
        """Stage times: returns (ms[8], calls[8]) accumulated since the last
        enable; enable=1/0 switches timing on/off and clears the totals."""
        ms = np.zeros(8)
        calls = np.zeros(8, np.int64)
        _check(self.L.pgk_profile(self.h, int(enable), ms.ctypes.data, calls.ctypes.data),
               "pgk_profile")
        return ms, calls

    def tune(self, knob, value):
        """Set a launch-shape choice (abi.TUNE_*); see pgk_tune."""
        _check(self.L.pgk_tune(self.h, int(knob), int(value)), "pgk_tune")

    def profile_stages(self, mask):
        """Stages timed while profiling is on: bit k = stage k (PAIRS = 0)."""
        _check(self.L.pgk_profile_stages(self.h, int(mask)), "pgk_profile_stages")

    def synchronize(self):
        _check(self.L.pgk_synchronize(self.h), "pgk_synchronize")


def md_run_batch(ctxs, nstep, nstep_book, rcut_books, nsample):
    """Step independent replicas together (pgk_md_run_batch): every context
    has had md_setup and md_set_velocities.  Returns (n, nstep//nsample + 1, 3)
    rows of kinetic, potential and total energy per replica."""
    n = len(ctxs)
    L = ctxs[0].L
    hs = (C.c_void_p * n)(*[c.h.value for c in ctxs])
    rb = np.ascontiguousarray(rcut_books, np.float64)
    out = np.zeros((n, nstep // nsample + 1, 3))
    _check(L.pgk_md_run_batch(C.cast(hs, C.c_void_p), n, int(nstep), int(nstep_book),
                              rb.ctypes.data, int(nsample), out.ctypes.data), "pgk_md_run_batch")
    return out
