"""ctypes mirror of include/pgk_force.h and marshalling of system dicts.

The structs here must match the C declarations field for field; the loader of
libpgkforce.so checks pgk_abi_version() against ABI_VERSION.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

ABI_VERSION = 1

PGK_OK, PGK_EINVAL, PGK_EDEVICE, PGK_ENOMEM, PGK_ESTATE, PGK_EOVERFLOW = 0, -1, -2, -3, -4, -5
PREC = {"fp64": 0, "mixed": 1, "fp32": 2}
MTS_LONG, MTS_MED, MTS_SHORT = 1, 2, 3
HF_VOLUME, HF_SURFACE, HF_BULK = 0, 1, 2
TUNE_PAIR_CHUNK_BELOW, TUNE_PME_OVERLAP, TUNE_PAIR_CAP_BLOCKS = 1, 2, 3
TUNE_DEFAULT_CHUNK_BELOW = 4352
HFTYP_ATOM, HFTYP_MOLE = 1, 2

ENE_NAMES = ["pot_tot", "pot_nonbon", "pot_vdw", "pot_elc", "pot_ewk", "pot_ewc",
             "pot_ewnq", "pot_vdw14", "pot_elc14", "pot_bond", "pot_angl", "pot_tors",
             "pot_torsrb", "pot_viri_coul", "pot_viri_lj", "atm_viri_bond",
             "atm_viri_angl", "atm_viri_tors", "atm_viri_14"]
NENE = len(ENE_NAMES)
TENS_NAMES = ["pot_virit_coul", "pot_virit_lj", "atm_virit_bond", "atm_virit_angl",
              "atm_virit_tors", "atm_virit_14"]
NTENS = len(TENS_NAMES)
HF_NAMES = ["bo", "an", "to", "14", "elin", "ljin", "elc", "vdw"]
NHF = len(HF_NAMES)
# calheatf output terms: hf[region, term, comp]
HFJ_NAMES = ["kin"] + ["pot_" + k for k in HF_NAMES] + ["viri_" + k for k in HF_NAMES] + ["tot"]
NHFJ = len(HFJ_NAMES)
HFJ_KIN, HFJ_POT, HFJ_VIRI, HFJ_TOT = 0, 1, 1 + NHF, 1 + 2 * NHF

_dp = C.POINTER(C.c_double)
_ip = C.POINTER(C.c_int32)


class SystemDesc(C.Structure):
    _fields_ = [
        ("natom", C.c_int32), ("ntype", C.c_int32),
        ("charge", _dp), ("atype", _ip), ("molid", _ip),
        ("lj_eps", _dp), ("lj_rmin", _dp), ("inttyp", _ip),
        ("nexcl", C.c_int32), ("excl_index", _ip), ("excl_list", _ip),
        ("nbond", C.c_int32), ("nbondtyp", C.c_int32),
        ("bond_ij", _ip), ("bond_type", _ip), ("bond_k", _dp), ("bond_r0", _dp),
        ("nangl", C.c_int32), ("nangltyp", C.c_int32),
        ("angl_ijk", _ip), ("angl_type", _ip), ("angl_k", _dp), ("angl_theta0", _dp),
        ("ntors", C.c_int32), ("ntorstyp", C.c_int32),
        ("tors_ijkl", _ip), ("tors_type", _ip), ("tors_ring14", _dp),
        ("tors_divfac", _ip), ("tors_barhig", _dp), ("tors_phsang", _dp),
        ("tors_period", _dp), ("tors_divfac_vdw", _dp), ("tors_divfac_elc", _dp),
        ("tors_if14", _ip),
        ("ntorsrb", C.c_int32), ("ntorsrbtyp", C.c_int32),
        ("torsrb_ijkl", _ip), ("torsrb_type", _ip), ("torsrb_ring14", _dp),
        ("torsrb_c", _dp), ("torsrb_divfac_vdw", _dp), ("torsrb_divfac_elc", _dp),
        ("torsrb_if14", _ip),
        ("alpha", C.c_double), ("rrcut", C.c_double), ("rcut", C.c_double),
        ("nspltbl", C.c_int32), ("nfft1", C.c_int32), ("nfft2", C.c_int32),
        ("nfft3", C.c_int32), ("pme_order", C.c_int32), ("ifnetqcorrp", C.c_int32),
        ("box", _dp),
    ]


class Mts(C.Structure):
    _fields_ = [("bond", C.c_int32), ("angl", C.c_int32), ("tors", C.c_int32),
                ("torsrb", C.c_int32), ("vdw", C.c_int32), ("ewk", C.c_int32),
                ("nonb14", C.c_int32)]


class ForceOut(C.Structure):
    _fields_ = [("force", _dp), ("ene", _dp), ("virt", _dp),
                ("for_viri_coul", _dp), ("for_viri_lj", _dp)]


class HfDesc(C.Structure):
    _fields_ = [("mode", C.c_int32), ("nregion", C.c_int32), ("z1", _dp), ("z2", _dp),
                ("hftyp", _ip), ("nmol", C.c_int32), ("molecom", _dp)]


class HfOut(C.Structure):
    _fields_ = [("pot_atm", _dp), ("viri_atm", _dp)]


class MdDesc(C.Structure):
    _fields_ = [("mass", _dp), ("ncons", C.c_int32), ("cons_ij", _ip), ("cons_d", _dp),
                ("ngroup", C.c_int32), ("group_start", _ip), ("dt", C.c_double),
                ("tol", C.c_double)]


class NhcDesc(C.Structure):
    _fields_ = [("mchain", C.c_int32), ("temperature", C.c_double), ("tfreq", C.c_double),
                ("nfree", C.c_double), ("nyosh", C.c_int32), ("next", C.c_int32),
                ("xlogs", _dp), ("vlogs", _dp)]


LT_MAXREG = 16


class LangevinDesc(C.Structure):
    _fields_ = [("nregion", C.c_int32), ("x1", _dp), ("x2", _dp), ("y1", _dp), ("y2", _dp),
                ("z1", _dp), ("z2", _dp), ("temperature", _dp), ("damp", _dp),
                ("mol_com", _ip), ("seed", C.c_uint64), ("nstream", C.c_int32),
                ("strm_z1", _dp), ("strm_z2", _dp), ("strm_vel", _dp)]


def dptr(a):
    return None if a is None else a.ctypes.data_as(_dp)


def iptr(a):
    return None if a is None else a.ctypes.data_as(_ip)


def _f64(a):
    return np.ascontiguousarray(a, dtype=np.float64)


def _i32(a):
    return np.ascontiguousarray(a, dtype=np.int32)


def make_desc(sysd, excl=None, ifnetqcorrp=0):
    """Build a SystemDesc from a system dict.  Returns (desc, keepalive) where
    keepalive holds every array the desc points into."""
    from .systems import exclusions
    keep = {}

    def f(name, arr):
        keep[name] = _f64(arr)
        return dptr(keep[name])

    def i(name, arr):
        keep[name] = _i32(arr)
        return iptr(keep[name])

    if excl is None:
        excl = exclusions(sysd)
    eidx, elst = excl
    d = SystemDesc()
    d.natom = int(sysd["natom"])
    d.ntype = int(sysd["lj_eps"].shape[0])
    d.charge = f("charge", sysd["charge"])
    d.atype = i("atype", sysd["atype"])
    d.molid = i("molid", sysd["molid"])
    # column-major ntype x ntype tables (Fortran vdw_welij(i,j))
    d.lj_eps = f("lj_eps", np.asfortranarray(sysd["lj_eps"]).ravel(order="F"))
    d.lj_rmin = f("lj_rmin", np.asfortranarray(sysd["lj_rmin"]).ravel(order="F"))
    d.inttyp = i("inttyp", np.asfortranarray(sysd["inttyp"]).ravel(order="F"))
    d.nexcl = int(len(elst))
    d.excl_index = i("excl_index", eidx)
    d.excl_list = i("excl_list", elst)
    d.nbond = int(len(sysd["bond_type"]))
    d.nbondtyp = int(len(sysd["bond_k"]))
    d.bond_ij = i("bond_ij", sysd["bond_ij"].reshape(-1))
    d.bond_type = i("bond_type", sysd["bond_type"])
    d.bond_k = f("bond_k", sysd["bond_k"])
    d.bond_r0 = f("bond_r0", sysd["bond_r0"])
    d.nangl = int(len(sysd["angl_type"]))
    d.nangltyp = int(len(sysd["angl_k"]))
    d.angl_ijk = i("angl_ijk", sysd["angl_ijk"].reshape(-1))
    d.angl_type = i("angl_type", sysd["angl_type"])
    d.angl_k = f("angl_k", sysd["angl_k"])
    d.angl_theta0 = f("angl_theta0", sysd["angl_theta0"])
    d.ntors = int(len(sysd["tors_type"]))
    d.ntorstyp = int(len(sysd["tors_barhig"]))
    d.tors_ijkl = i("tors_ijkl", sysd["tors_ijkl"].reshape(-1))
    d.tors_type = i("tors_type", sysd["tors_type"])
    d.tors_ring14 = f("tors_ring14", sysd["tors_ring14"])
    d.tors_divfac = i("tors_divfac", sysd["tors_divfac"])
    d.tors_barhig = f("tors_barhig", sysd["tors_barhig"])
    d.tors_phsang = f("tors_phsang", sysd["tors_phsang"])
    d.tors_period = f("tors_period", sysd["tors_period"])
    d.tors_divfac_vdw = f("tors_divfac_vdw", sysd["tors_divfac_vdw"])
    d.tors_divfac_elc = f("tors_divfac_elc", sysd["tors_divfac_elc"])
    d.tors_if14 = i("tors_if14", sysd["tors_if14"])
    d.ntorsrb = int(len(sysd["torsrb_type"]))
    d.ntorsrbtyp = int(sysd["torsrb_c"].shape[0])
    d.torsrb_ijkl = i("torsrb_ijkl", sysd["torsrb_ijkl"].reshape(-1))
    d.torsrb_type = i("torsrb_type", sysd["torsrb_type"])
    d.torsrb_ring14 = f("torsrb_ring14", sysd["torsrb_ring14"])
    d.torsrb_c = f("torsrb_c", sysd["torsrb_c"].reshape(-1))
    d.torsrb_divfac_vdw = f("torsrb_divfac_vdw", sysd["torsrb_divfac_vdw"])
    d.torsrb_divfac_elc = f("torsrb_divfac_elc", sysd["torsrb_divfac_elc"])
    d.torsrb_if14 = i("torsrb_if14", sysd["torsrb_if14"])
    d.alpha = float(sysd["alpha"])
    d.rrcut = float(sysd["rrcut"])
    d.rcut = float(sysd["rcut"])
    d.nspltbl = int(sysd["nspltbl"])
    nf = [int(x) for x in sysd["nfft"]]
    d.nfft1, d.nfft2, d.nfft3 = nf
    d.pme_order = int(sysd["pme_order"])
    d.ifnetqcorrp = int(ifnetqcorrp)
    d.box = f("box", sysd["box"])
    return d, keep


def mts_all(level=MTS_LONG):
    """Every term on one MTS level (a single calforce call evaluates all)."""
    return Mts(level, level, level, level, level, level, level)
