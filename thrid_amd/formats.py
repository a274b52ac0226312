"""The reference's on-disk formats through libpgkforce.so (host code, no GPU):
the wrsta / rdstarec state file, outene's energy rows and the .ini control
script.  State-file values stay in the file's SI units (m, m/s, 1/s, Pa)."""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import pgk


class StaHeader(C.Structure):
    _fields_ = [("npoly", C.c_int32), ("nwater", C.c_int32), ("nmatom", C.c_int32),
                ("natom", C.c_int32), ("mchain", C.c_int32), ("nat_poly", C.c_int32),
                ("nat_water", C.c_int32), ("nat_matom", C.c_int32),
                ("box", C.c_double * 3), ("nstep", C.c_int64)]


def _lib():
    L = pgk.load_library()
    vp = C.c_void_p
    L.pgk_sta_read_header.argtypes = [C.c_char_p, C.POINTER(StaHeader)]
    L.pgk_sta_read.argtypes = [C.c_char_p, C.POINTER(StaHeader)] + [vp] * 6
    L.pgk_sta_write.argtypes = [C.c_char_p, C.POINTER(StaHeader)] + [vp] * 6
    L.pgk_out_ene_row.argtypes = [C.c_int32, vp, C.c_int32, C.c_int32, C.c_char_p, C.c_int32]
    L.pgk_out_pre_row.argtypes = [C.c_int32, vp, C.c_int32, C.c_int32, C.c_char_p, C.c_int32]
    L.pgk_out_pre_header.argtypes = [C.c_char_p, C.c_int32]
    L.pgk_out_htf_header.argtypes = [C.c_int32, C.c_int32, C.c_char_p, C.c_int32]
    L.pgk_out_htf_row.argtypes = [C.c_int32, vp, C.c_int32, C.c_int32, C.c_int32, C.c_double,
                                  C.c_char_p, C.c_int32]
    L.pgk_ini_read.argtypes = [C.c_char_p, C.POINTER(vp)]
    L.pgk_ini_free.argtypes = [vp]
    L.pgk_ini_free.restype = None
    L.pgk_ini_count.argtypes = [vp, C.c_char_p]
    L.pgk_ini_count.restype = C.c_int32
    L.pgk_ini_nvalues.argtypes = [vp, C.c_char_p, C.c_int32]
    L.pgk_ini_nvalues.restype = C.c_int32
    L.pgk_ini_get_str.argtypes = [vp, C.c_char_p, C.c_int32, C.c_int32, C.c_char_p, C.c_int32]
    L.pgk_ini_get_double.argtypes = [vp, C.c_char_p, C.c_int32, C.c_int32, vp]
    L.pgk_ini_get_int.argtypes = [vp, C.c_char_p, C.c_int32, C.c_int32, vp]
    L.pgk_ini_get_bool.argtypes = [vp, C.c_char_p, C.c_int32, C.c_int32, vp]
    for f in ("pgk_sta_read_header", "pgk_sta_read", "pgk_sta_write", "pgk_out_ene_row",
              "pgk_out_pre_row", "pgk_out_pre_header", "pgk_out_htf_header", "pgk_out_htf_row",
              "pgk_ini_read", "pgk_ini_get_str", "pgk_ini_get_double", "pgk_ini_get_int",
              "pgk_ini_get_bool"):
        getattr(L, f).restype = C.c_int
    return L


def read_sta(path):
    """State file -> dict (header fields, pos/vel (natom, 3), types, nhc
    (mchain, 2) = (xlogs, vlogs), andersen (2,), aniso (12,))."""
    L = _lib()
    h = StaHeader()
    pgk._check(L.pgk_sta_read_header(str(path).encode(), C.byref(h)), "pgk_sta_read_header")
    n, m = h.natom, h.mchain
    pos, vel = np.zeros((n, 3)), np.zeros((n, 3))
    types = C.create_string_buffer(4 * max(n, 1))
    nhc, ander, aniso = np.zeros((max(m, 1), 2)), np.zeros(2), np.zeros(12)
    pgk._check(L.pgk_sta_read(str(path).encode(), C.byref(h), pos.ctypes.data, vel.ctypes.data,
                              C.addressof(types), nhc.ctypes.data, ander.ctypes.data,
                              aniso.ctypes.data), "pgk_sta_read")
    raw = types.raw
    return dict(npoly=h.npoly, nwater=h.nwater, nmatom=h.nmatom, natom=n, mchain=m,
                sections=(h.nat_poly, h.nat_water, h.nat_matom), box=np.array(h.box[:]),
                nstep=h.nstep, pos=pos, vel=vel,
                types=[raw[4 * i:4 * i + 4].rstrip(b"\0").decode() for i in range(n)],
                nhc=nhc[:m], andersen=ander, aniso=aniso)


def write_sta(path, d):
    L = _lib()
    h = StaHeader()
    h.npoly, h.nwater, h.nmatom = d["npoly"], d["nwater"], d["nmatom"]
    h.natom, h.mchain = d["natom"], d["mchain"]
    h.nat_poly, h.nat_water, h.nat_matom = d["sections"]
    h.box[:] = list(d["box"])
    h.nstep = int(d["nstep"])
    pos = np.ascontiguousarray(d["pos"], np.float64)
    vel = np.ascontiguousarray(d["vel"], np.float64)
    types = b"".join(t.encode()[:3].ljust(4, b"\0") for t in d["types"])
    tb = C.create_string_buffer(types, max(len(types), 4))
    nhc = np.ascontiguousarray(d["nhc"], np.float64).reshape(-1, 2)
    if nhc.shape[0] == 0:
        nhc = np.zeros((1, 2))
    ander = np.ascontiguousarray(d["andersen"], np.float64)
    aniso = np.ascontiguousarray(d["aniso"], np.float64)
    pgk._check(L.pgk_sta_write(str(path).encode(), C.byref(h), pos.ctypes.data, vel.ctypes.data,
                               C.addressof(tb), nhc.ctypes.data, ander.ctypes.data,
                               aniso.ctypes.data), "pgk_sta_write")


def out_ene_row(step, vals, high=False):
    """One out_ene.dat data row (with its newline)."""
    L = _lib()
    v = np.ascontiguousarray(vals, np.float64)
    buf = C.create_string_buffer(32 * (len(v) + 2))
    pgk._check(L.pgk_out_ene_row(int(step), v.ctypes.data, len(v), int(bool(high)), buf,
                                 len(buf)), "pgk_out_ene_row")
    return buf.value.decode()


def out_pre_header():
    """The header line of out_pre.dat (with its newline)."""
    L = _lib()
    buf = C.create_string_buffer(32 * 84)
    pgk._check(L.pgk_out_pre_header(buf, len(buf)), "pgk_out_pre_header")
    return buf.value.decode()


def out_pre_row(step, vals, high=False):
    """One out_pre.dat data row (with its newline): the 81 values in the
    file's order."""
    L = _lib()
    v = np.ascontiguousarray(vals, np.float64)
    buf = C.create_string_buffer(32 * (len(v) + 2))
    pgk._check(L.pgk_out_pre_row(int(step), v.ctypes.data, len(v), int(bool(high)), buf,
                                 len(buf)), "pgk_out_pre_row")
    return buf.value.decode()


def out_htf_header(nregion, all_dir=True):
    """The header line of out_htf.dat for nregion regions."""
    L = _lib()
    buf = C.create_string_buffer(32 * (36 * 3 * int(nregion) + 2))
    pgk._check(L.pgk_out_htf_header(int(nregion), int(bool(all_dir)), buf, len(buf)),
               "pgk_out_htf_header")
    return buf.value.decode()


def out_htf_row(step, hf, jref=1.0, all_dir=True, high=False):
    """One out_htf.dat data row from hf [nregion, NHFJ, 3] (calheatf's
    layout) times jref."""
    L = _lib()
    h = np.ascontiguousarray(hf, np.float64)
    nreg = h.shape[0]
    buf = C.create_string_buffer(32 * (36 * 3 * nreg + 2))
    pgk._check(L.pgk_out_htf_row(int(step), h.ctypes.data, nreg, int(bool(all_dir)),
                                 int(bool(high)), float(jref), buf, len(buf)), "pgk_out_htf_row")
    return buf.value.decode()


class Ini:
    """The words of a .ini control script (rdscript.F90 / rdfree.F90)."""

    def __init__(self, path):
        self.L = _lib()
        self.h = C.c_void_p()
        pgk._check(self.L.pgk_ini_read(str(path).encode(), C.byref(self.h)), "pgk_ini_read")

    def close(self):
        if self.h:
            self.L.pgk_ini_free(self.h)
            self.h = C.c_void_p()

    def __del__(self):
        self.close()

    def count(self, key):
        return self.L.pgk_ini_count(self.h, key.encode())

    def nvalues(self, key, occ=0):
        return self.L.pgk_ini_nvalues(self.h, key.encode(), occ)

    def str(self, key, idx=0, occ=0):
        buf = C.create_string_buffer(256)
        pgk._check(self.L.pgk_ini_get_str(self.h, key.encode(), occ, idx, buf, 256),
                   "pgk_ini_get_str")
        return buf.value.decode()

    def float(self, key, idx=0, occ=0):
        v = C.c_double()
        pgk._check(self.L.pgk_ini_get_double(self.h, key.encode(), occ, idx, C.byref(v)),
                   "pgk_ini_get_double")
        return v.value

    def int(self, key, idx=0, occ=0):
        v = C.c_int64()
        pgk._check(self.L.pgk_ini_get_int(self.h, key.encode(), occ, idx, C.byref(v)),
                   "pgk_ini_get_int")
        return v.value

    def bool(self, key, idx=0, occ=0):
        v = C.c_int32()
        pgk._check(self.L.pgk_ini_get_bool(self.h, key.encode(), occ, idx, C.byref(v)),
                   "pgk_ini_get_bool")
        return bool(v.value)
