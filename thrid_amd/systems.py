"""Synthetic peachgk_md systems in the reference's non-dimensional units.

Units follow the reference (main_md.F90:388-396, calbase.F90:103-120):
length 1 A, energy eref = 6.9511042e-21 J, charge already multiplied by
1/sqrt(4*pi*eps0) (prepmd.F90:239-242), LJ written as E*[(R/r)^12 - 2(R/r)^6]
with R = R_min = 2^(1/6) sigma (ewaldr.F90:21), epsilon_ij = sqrt(ei*ej) and
R_ij arithmetic (rdpara.F90:551-555), bond energy K(r-r0)^2 and angle energy
K(theta-theta0)^2 with no factor 1/2 (calbond.F90:85-86, calangl.F90).

Every system is a plain dict of numpy arrays (1-based atom indices in all
topology arrays, as the reference stores them) so it can be handed to the GPU
library, to the oracle and to the compiled reference without conversion.
"""
from __future__ import annotations

import math
from dataclasses import dataclass

import numpy as np

XREF = 1.0e-10
EREF = 6.9511042e-21
QREF = 1.602117733e-19
EPS0 = 8.854187817e-12
KCAL_MOL_J = 4184.0 / 6.0221367e23
# charge scale: q[e] -> q[non-d] / sqrt(4 pi eps0[non-d])
QSCALE = math.sqrt(QREF ** 2 / (4.0 * math.pi * EPS0 * EREF * XREF))
RMIN_PER_SIGMA = 2.0 ** (1.0 / 6.0)
DEG = math.pi / 180.0


def kcal(e: float) -> float:
    """kcal/mol -> reference energy unit."""
    return e * KCAL_MOL_J / EREF


# ---------------------------------------------------------------------------
# force-field tables (synthetic, in the ranges the workloads name)
# ---------------------------------------------------------------------------
@dataclass(frozen=True)
class AtomType:
    name: str
    sigma: float      # A
    eps: float        # kcal/mol
    mass: float       # g/mol


ATOM_TYPES = [
    AtomType("OB", 3.166, 0.1553, 15.9994),   # SPC/E oxygen
    AtomType("HB", 0.0, 0.0, 1.008),          # SPC/E hydrogen (no LJ)
    AtomType("CH", 3.50, 0.066, 12.011),      # alcohol alpha carbon
    AtomType("CT", 3.50, 0.066, 12.011),      # aliphatic carbon
    AtomType("OH", 3.07, 0.17, 15.9994),      # hydroxyl oxygen
    AtomType("HO", 2.50, 0.03, 1.008),        # hydroxyl hydrogen
    AtomType("H1", 2.50, 0.03, 1.008),        # H on alpha carbon
    AtomType("HC", 2.50, 0.03, 1.008),        # aliphatic H
    AtomType("NA", 2.35, 0.13, 22.98977),     # Na+
    AtomType("CL", 4.40, 0.10, 35.453),       # Cl-
]
TYPE_INDEX = {t.name: i + 1 for i, t in enumerate(ATOM_TYPES)}   # 1-based


def lj_tables():
    """vdw_welij / vdw_radij / inter_inttyp over all atom types (column-major)."""
    nt = len(ATOM_TYPES)
    eps = np.array([kcal(t.eps) for t in ATOM_TYPES])
    rmin = np.array([t.sigma * RMIN_PER_SIGMA for t in ATOM_TYPES])
    welij = np.sqrt(np.outer(eps, eps))
    radij = 0.5 * (rmin[:, None] + rmin[None, :])
    inttyp = np.ones((nt, nt), dtype=np.int32)   # INTTYPE_VDW everywhere
    return welij, radij, inttyp


# bond types: (name, K [kcal/mol/A^2] for E = K(r-r0)^2, r0 [A])
BOND_TYPES = [
    ("HB-OB", 0.0, 1.0),        # rigid water: K = 0, kept for exclusions
    ("HB-HB", 0.0, 1.63299),
    ("CT-CH", 310.0, 1.526),
    ("CH-OH", 320.0, 1.410),
    ("CH-H1", 340.0, 1.090),
    ("CT-CT", 310.0, 1.526),
    ("CT-HC", 340.0, 1.090),
    ("HO-OH", 553.0, 0.960),
]
BOND_INDEX = {n: i + 1 for i, (n, _, _) in enumerate(BOND_TYPES)}

# angle types: (name, K [kcal/mol/rad^2] for E = K(th-th0)^2, th0 [deg])
ANGLE_TYPES = [
    ("HB-OB-HB", 0.0, 109.47),
    ("CH-CT-CT", 40.0, 109.5),
    ("CH-CT-HC", 50.0, 109.5),
    ("CH-OH-HO", 55.0, 108.5),
    ("CT-CH-OH", 50.0, 109.5),
    ("CT-CH-H1", 50.0, 109.5),
    ("CT-CT-HC", 50.0, 109.5),
    ("H1-CH-OH", 50.0, 109.5),
    ("H1-CH-H1", 35.0, 109.5),
    ("HC-CT-HC", 35.0, 109.5),
    ("CT-CT-CT", 40.0, 109.5),
]
ANGLE_INDEX = {n: i + 1 for i, (n, _, _) in enumerate(ANGLE_TYPES)}

# Ryckaert-Bellemans torsion types: (name, C0..C5 [kcal/mol])
RB_TYPES = [
    ("X-CT-CT-X", (0.30, 0.90, 0.00, -1.20, 0.00, 0.00)),
    ("X-CH-CT-X", (0.28, 0.84, 0.05, -1.12, 0.02, -0.01)),
    ("X-CH-OH-X", (0.22, 0.66, 0.10, -0.88, 0.03, 0.02)),
    ("C-CT-CT-C", (2.22, 2.91, -0.73, -4.40, 0.00, 0.00)),
]
RB_INDEX = {n: i + 1 for i, (n, _) in enumerate(RB_TYPES)}

# periodic torsion types: (name, divfac, barhig [kcal/mol], phase [deg], n)
TORS_TYPES = [
    ("X-CT-CT-X", 9, 1.40, 0.0, 3.0),
    ("X-CH-CT-X", 9, 1.40, 0.0, 3.0),
    ("X-CH-OH-X", 3, 0.50, 0.0, 3.0),
    ("C-C-C-C", 1, 0.25, 180.0, 2.0),
    ("C-C-C-O", 1, 0.20, 0.0, 1.0),
]
TORS_INDEX = {n: i + 1 for i, (n, *_r) in enumerate(TORS_TYPES)}

# ---------------------------------------------------------------------------
# molecule templates: local geometry, types, charges, topology (1-based local)
# ---------------------------------------------------------------------------
_TH_W = 109.47 * DEG


def _water_template():
    # atom order H, O, H as in the reference water topology (H2O_topOB.dat)
    oh = 1.0
    pos = np.array([
        [oh * math.sin(_TH_W / 2), 0.0, oh * math.cos(_TH_W / 2)],
        [0.0, 0.0, 0.0],
        [-oh * math.sin(_TH_W / 2), 0.0, oh * math.cos(_TH_W / 2)],
    ])
    return dict(
        pos=pos,
        types=["HB", "OB", "HB"],
        charges=[0.4238, -0.8476, 0.4238],
        bonds=[(1, 2, "HB-OB"), (1, 3, "HB-HB"), (2, 3, "HB-OB")],
        angles=[(1, 2, 3, "HB-OB-HB")],
        tors=[],
    )


def _place(parent, ref, bond, angle, dihedral, ref2=None):
    """z-matrix style placement of a new atom bonded to `parent`."""
    a = np.asarray(parent, float)
    b = np.asarray(ref, float)
    c = np.asarray(ref2 if ref2 is not None else b + np.array([0.3, 0.7, 0.1]), float)
    bc = b - c
    bc /= np.linalg.norm(bc)
    ab = a - b
    ab /= np.linalg.norm(ab)
    n = np.cross(bc, ab)
    n /= np.linalg.norm(n)
    m = np.cross(n, ab)
    d = np.array([-bond * math.cos(angle),
                  bond * math.sin(angle) * math.cos(dihedral),
                  bond * math.sin(angle) * math.sin(dihedral)])
    return a + d[0] * ab + d[1] * m + d[2] * n


def _propanol_template():
    # numbering of the reference's C3H7OH topology (C3H7OH_top.dat):
    # 1 CH, 2 CT, 3 CT, 4 OH, 5-6 H1, 7-8 HC, 9-11 HC, 12 HO
    th = 109.5 * DEG
    p = [None] * 12
    p[0] = np.zeros(3)
    p[1] = np.array([1.526, 0.0, 0.0])
    p[2] = _place(p[1], p[0], 1.526, th, 0.0, np.array([0.0, 1.0, 0.0]))
    p[3] = _place(p[0], p[1], 1.410, th, 180.0 * DEG, p[2])
    p[4] = _place(p[0], p[1], 1.090, th, 60.0 * DEG, p[2])
    p[5] = _place(p[0], p[1], 1.090, th, -60.0 * DEG, p[2])
    p[6] = _place(p[1], p[2], 1.090, th, 120.0 * DEG, p[0])
    p[7] = _place(p[1], p[2], 1.090, th, -120.0 * DEG, p[0])
    p[8] = _place(p[2], p[1], 1.090, th, 180.0 * DEG, p[0])
    p[9] = _place(p[2], p[1], 1.090, th, 60.0 * DEG, p[0])
    p[10] = _place(p[2], p[1], 1.090, th, -60.0 * DEG, p[0])
    p[11] = _place(p[3], p[0], 0.960, 108.5 * DEG, 180.0 * DEG, p[1])
    pos = np.array(p)
    types = ["CH", "CT", "CT", "OH", "H1", "H1", "HC", "HC", "HC", "HC", "HC", "HO"]
    charges = [0.06, -0.04, -0.18, -0.68, 0.06, 0.06, 0.06, 0.06, 0.06, 0.06, 0.06, 0.42]
    bonds = [(1, 2, "CT-CH"), (1, 4, "CH-OH"), (1, 5, "CH-H1"), (1, 6, "CH-H1"),
             (2, 3, "CT-CT"), (2, 7, "CT-HC"), (2, 8, "CT-HC"), (3, 9, "CT-HC"),
             (3, 10, "CT-HC"), (3, 11, "CT-HC"), (4, 12, "HO-OH")]
    angles = [(1, 2, 3, "CH-CT-CT"), (1, 2, 7, "CH-CT-HC"), (1, 2, 8, "CH-CT-HC"),
              (1, 4, 12, "CH-OH-HO"), (2, 1, 4, "CT-CH-OH"), (2, 1, 5, "CT-CH-H1"),
              (2, 1, 6, "CT-CH-H1"), (2, 3, 9, "CT-CT-HC"), (2, 3, 10, "CT-CT-HC"),
              (2, 3, 11, "CT-CT-HC"), (3, 2, 7, "CT-CT-HC"), (3, 2, 8, "CT-CT-HC"),
              (4, 1, 5, "H1-CH-OH"), (4, 1, 6, "H1-CH-OH"), (5, 1, 6, "H1-CH-H1"),
              (7, 2, 8, "HC-CT-HC"), (9, 3, 10, "HC-CT-HC"), (9, 3, 11, "HC-CT-HC"),
              (10, 3, 11, "HC-CT-HC")]
    tors = [(1, 2, 3, 9, "X-CT-CT-X"), (1, 2, 3, 10, "X-CT-CT-X"), (1, 2, 3, 11, "X-CT-CT-X"),
            (2, 1, 4, 12, "X-CH-OH-X"), (3, 2, 1, 4, "X-CH-CT-X"), (3, 2, 1, 5, "X-CH-CT-X"),
            (3, 2, 1, 6, "X-CH-CT-X"), (4, 1, 2, 7, "X-CH-CT-X"), (4, 1, 2, 8, "X-CH-CT-X"),
            (5, 1, 2, 7, "X-CH-CT-X"), (5, 1, 2, 8, "X-CH-CT-X"), (5, 1, 4, 12, "X-CH-OH-X"),
            (6, 1, 2, 7, "X-CH-CT-X"), (6, 1, 2, 8, "X-CH-CT-X"), (6, 1, 4, 12, "X-CH-OH-X"),
            (7, 2, 3, 9, "X-CT-CT-X"), (7, 2, 3, 10, "X-CT-CT-X"), (7, 2, 3, 11, "X-CT-CT-X"),
            (8, 2, 3, 9, "X-CT-CT-X"), (8, 2, 3, 10, "X-CT-CT-X"), (8, 2, 3, 11, "X-CT-CT-X")]
    return dict(pos=pos, types=types, charges=charges, bonds=bonds, angles=angles, tors=tors)


def _butane_template():
    # all-atom n-butane: 1-4 carbons, 5-14 hydrogens
    th = 109.5 * DEG
    c = [np.zeros(3), np.array([1.526, 0, 0])]
    c.append(_place(c[1], c[0], 1.526, th, 0.0, np.array([0.0, 1.0, 0.0])))
    c.append(_place(c[2], c[1], 1.526, th, 180.0 * DEG, c[0]))
    h = []
    h.append(_place(c[0], c[1], 1.09, th, 180.0 * DEG, c[2]))
    h.append(_place(c[0], c[1], 1.09, th, 60.0 * DEG, c[2]))
    h.append(_place(c[0], c[1], 1.09, th, -60.0 * DEG, c[2]))
    h.append(_place(c[1], c[2], 1.09, th, 120.0 * DEG, c[0]))
    h.append(_place(c[1], c[2], 1.09, th, -120.0 * DEG, c[0]))
    h.append(_place(c[2], c[1], 1.09, th, 120.0 * DEG, c[3]))
    h.append(_place(c[2], c[1], 1.09, th, -120.0 * DEG, c[3]))
    h.append(_place(c[3], c[2], 1.09, th, 180.0 * DEG, c[1]))
    h.append(_place(c[3], c[2], 1.09, th, 60.0 * DEG, c[1]))
    h.append(_place(c[3], c[2], 1.09, th, -60.0 * DEG, c[1]))
    pos = np.array(c + h)
    types = ["CT"] * 4 + ["HC"] * 10
    charges = [-0.18, -0.12, -0.12, -0.18] + [0.06] * 10
    hmap = {1: [5, 6, 7], 2: [8, 9], 3: [10, 11], 4: [12, 13, 14]}
    bonds = [(1, 2, "CT-CT"), (2, 3, "CT-CT"), (3, 4, "CT-CT")]
    for ci, hs in hmap.items():
        bonds += [(ci, hh, "CT-HC") for hh in hs]
    angles = [(1, 2, 3, "CT-CT-CT"), (2, 3, 4, "CT-CT-CT")]
    nbr = {1: [2], 2: [1, 3], 3: [2, 4], 4: [3]}
    for ci, hs in hmap.items():
        for a in range(len(hs)):
            for b in range(a + 1, len(hs)):
                angles.append((hs[a], ci, hs[b], "HC-CT-HC"))
        for hh in hs:
            for cn in nbr[ci]:
                angles.append((hh, ci, cn, "CT-CT-HC"))
    tors = [(1, 2, 3, 4, "C-CT-CT-C")]
    for (j, k) in ((1, 2), (2, 3), (3, 4)):
        iset = hmap[j] + [x for x in nbr[j] if x != k]
        lset = hmap[k] + [x for x in nbr[k] if x != j]
        for i in iset:
            for l in lset:
                if (i, l) == (1, 4) or (i, l) == (4, 1):
                    continue
                tors.append((i, j, k, l, "X-CT-CT-X"))
    return dict(pos=pos, types=types, charges=charges, bonds=bonds, angles=angles, tors=tors)


TEMPLATES = {"water": _water_template, "propanol": _propanol_template,
             "butane": _butane_template}


def _random_rotation(rng):
    q = rng.normal(size=4)
    q /= np.linalg.norm(q)
    a, b, c, d = q
    return np.array([
        [a * a + b * b - c * c - d * d, 2 * (b * c - a * d), 2 * (b * d + a * c)],
        [2 * (b * c + a * d), a * a - b * b + c * c - d * d, 2 * (c * d - a * b)],
        [2 * (b * d - a * c), 2 * (c * d + a * b), a * a - b * b - c * c + d * d],
    ])


class _Builder:
    def __init__(self):
        self.pos, self.types, self.q, self.mol = [], [], [], []
        self.bonds, self.angles, self.tors = [], [], []
        self.nmol = 0
        self.natom = 0

    def add_checked(self, tmpl, centre, rng, box, dmin=2.2, tries=200):
        """Add with a random orientation that keeps every atom at least `dmin`
        (minimum image) from the atoms already placed."""
        placed = np.concatenate(self.pos, axis=0) if self.pos else np.zeros((0, 3))
        box = np.asarray(box, float)
        local = tmpl["pos"] - tmpl["pos"].mean(axis=0)
        best, bestd = None, -1.0
        near = np.zeros((0, 3))
        if placed.shape[0]:
            d0 = placed - centre
            d0 -= box * np.round(d0 / box)
            near = d0[np.all(np.abs(d0) < 8.0, axis=1)]
        for _ in range(tries):
            rot = _random_rotation(rng)
            if near.shape[0] == 0:
                best = rot
                break
            dr = (local @ rot.T)[:, None, :] - near[None, :, :]
            dm = np.sqrt((dr ** 2).sum(-1)).min()
            if dm > bestd:
                best, bestd = rot, dm
            if dm >= dmin:
                break
        self.add(tmpl, centre, best)

    def add(self, tmpl, centre, rot):
        base = self.natom
        self.nmol += 1
        local = tmpl["pos"] - tmpl["pos"].mean(axis=0)
        self.pos.append(local @ rot.T + centre)
        self.types += [TYPE_INDEX[t] for t in tmpl["types"]]
        self.q += list(tmpl["charges"])
        self.mol += [self.nmol] * len(tmpl["types"])
        self.bonds += [(base + i, base + j, BOND_INDEX[n]) for i, j, n in tmpl["bonds"]]
        self.angles += [(base + i, base + j, base + k, ANGLE_INDEX[n])
                        for i, j, k, n in tmpl["angles"]]
        self.tors += [(base + i, base + j, base + k, base + l, n)
                      for i, j, k, l, n in tmpl["tors"]]
        self.natom += len(tmpl["types"])

    def add_ion(self, name, charge, centre):
        self.nmol += 1
        self.pos.append(np.asarray(centre, float)[None, :])
        self.types.append(TYPE_INDEX[name])
        self.q.append(charge)
        self.mol.append(self.nmol)
        self.natom += 1


def _finish(b: _Builder, box, *, alpha, rrcut, rcut, rcut_book, nfft, order,
            torsion_form="rb", nspltbl=None, name="", seed=0):
    welij, radij, inttyp = lj_tables()
    pos = np.concatenate(b.pos, axis=0) if b.pos else np.zeros((0, 3))
    n = pos.shape[0]
    sysd = dict(
        name=name, seed=seed,
        natom=n, box=np.array(box, float),
        pos=np.ascontiguousarray(pos, dtype=np.float64),
        charge=np.array(b.q, float) * QSCALE,
        atype=np.array(b.types, np.int32), molid=np.array(b.mol, np.int32),
        nmol=b.nmol,
        mass=np.array([ATOM_TYPES[t - 1].mass for t in b.types], float),
        lj_eps=welij, lj_rmin=radij, inttyp=inttyp,
        alpha=float(alpha), rrcut=float(rrcut), rcut=float(rcut),
        rcut_book=float(rcut_book),
        nspltbl=int(nspltbl if nspltbl else max(100, int(round((rrcut + 2.0) / 0.01)))),
        nfft=np.array(nfft, np.int32), pme_order=int(order),
    )
    nb = len(b.bonds)
    sysd["bond_ij"] = np.array([(i, j) for i, j, _ in b.bonds], np.int32).reshape(nb, 2)
    sysd["bond_type"] = np.array([t for *_x, t in b.bonds], np.int32)
    sysd["bond_k"] = np.array([kcal(k) for _n, k, _r in BOND_TYPES])
    sysd["bond_r0"] = np.array([r for _n, _k, r in BOND_TYPES])
    na = len(b.angles)
    sysd["angl_ijk"] = np.array([a[:3] for a in b.angles], np.int32).reshape(na, 3)
    sysd["angl_type"] = np.array([a[3] for a in b.angles], np.int32)
    sysd["angl_k"] = np.array([kcal(k) for _n, k, _t in ANGLE_TYPES])
    sysd["angl_theta0"] = np.array([t * DEG for _n, _k, t in ANGLE_TYPES])
    rb = [t for t in b.tors] if torsion_form == "rb" else []
    per = [t for t in b.tors] if torsion_form == "periodic" else []
    nrb = len(rb)
    sysd["torsrb_ijkl"] = np.array([t[:4] for t in rb], np.int32).reshape(nrb, 4)
    sysd["torsrb_type"] = np.array([RB_INDEX[t[4]] if t[4] in RB_INDEX else 1 for t in rb], np.int32)
    sysd["torsrb_ring14"] = np.ones(nrb)
    sysd["torsrb_c"] = np.array([[kcal(c) for c in cs] for _n, cs in RB_TYPES]).reshape(-1, 6)
    sysd["torsrb_divfac_vdw"] = np.full(len(RB_TYPES), 2.0)
    sysd["torsrb_divfac_elc"] = np.full(len(RB_TYPES), 2.0)
    sysd["torsrb_if14"] = np.ones(len(RB_TYPES), np.int32)
    npt = len(per)
    tmap = {"X-CT-CT-X": "X-CT-CT-X", "X-CH-CT-X": "X-CH-CT-X", "X-CH-OH-X": "X-CH-OH-X",
            "C-CT-CT-C": "C-C-C-C"}
    sysd["tors_ijkl"] = np.array([t[:4] for t in per], np.int32).reshape(npt, 4)
    sysd["tors_type"] = np.array([TORS_INDEX[tmap.get(t[4], "C-C-C-O")] for t in per], np.int32)
    sysd["tors_ring14"] = np.ones(npt)
    sysd["tors_divfac"] = np.array([d for _n, d, *_r in TORS_TYPES], np.int32)
    sysd["tors_barhig"] = np.array([kcal(h) for _n, _d, h, _p, _m in TORS_TYPES])
    sysd["tors_phsang"] = np.array([p * DEG for _n, _d, _h, p, _m in TORS_TYPES])
    sysd["tors_period"] = np.array([m for *_r, m in TORS_TYPES], float)
    sysd["tors_divfac_vdw"] = np.full(len(TORS_TYPES), 2.0)
    sysd["tors_divfac_elc"] = np.full(len(TORS_TYPES), 2.0)
    sysd["tors_if14"] = np.ones(len(TORS_TYPES), np.int32)
    return sysd


def _lattice_sites(nx, ny, nz, lo, hi, rng, jitter):
    lo = np.asarray(lo, float)
    hi = np.asarray(hi, float)
    step = (hi - lo) / np.array([nx, ny, nz])
    g = np.stack(np.meshgrid(np.arange(nx), np.arange(ny), np.arange(nz), indexing="ij"),
                 -1).reshape(-1, 3)
    pts = lo + (g + 0.5) * step
    return pts + rng.uniform(-jitter, jitter, size=pts.shape)


def water_box(nside=6, box=18.64, *, seed=1, alpha=0.35, rcut=9.0, rcut_book=9.3,
              nfft=(24, 24, 24), order=4, nmol=None, ions=0, name="water"):
    """216 SPC/E-like rigid waters on a jittered lattice with random orientations
    (the 648-atom workload); with `nmol`/`ions` the same recipe fills bigger boxes."""
    rng = np.random.default_rng(seed)
    b = _Builder()
    tmpl = _water_template()
    nmol = nmol if nmol is not None else nside ** 3
    nsite = int(math.ceil((nmol + ions) ** (1.0 / 3.0) - 1e-9))
    sites = _lattice_sites(nsite, nsite, nsite, (0, 0, 0), (box, box, box), rng,
                           0.25 * min(1.0, box / nsite / 3.107))
    order_idx = rng.permutation(len(sites))
    sites = sites[order_idx]
    ion_sites = sites[:ions]
    for c in sites[ions:ions + nmol]:
        b.add(tmpl, c, _random_rotation(rng))
    for k, c in enumerate(ion_sites):
        if k % 2 == 0:
            b.add_ion("NA", 1.0, c)
        else:
            b.add_ion("CL", -1.0, c)
    return _finish(b, (box, box, box), alpha=alpha, rrcut=rcut, rcut=rcut,
                   rcut_book=rcut_book, nfft=nfft, order=order, name=name, seed=seed)


def water_ion_box_100k(seed=3):
    """100,001 atoms: 33,333 waters plus one Na+/Cl- pair in a 100.1 A box,
    SPME 128^3 order 4, alpha 0.32/A, 10 A cutoff (list skin 1 A)."""
    return water_box(box=100.1, nmol=33333, ions=2, seed=seed, alpha=0.32, rcut=10.0,
                     rcut_book=11.0, nfft=(128, 128, 128), order=4, name="water100k")


def propanol_water_slab(nwater=4096, nprop=512, box=(50.0, 50.0, 150.0), *, seed=2,
                        alpha=0.26, rcut=12.0, rcut_book=13.0, nfft=(64, 64, 192),
                        order=4, torsion_form="rb", name="slab"):
    """Central water slab with a 1-propanol layer on each face, vacuum beyond."""
    rng = np.random.default_rng(seed)
    b = _Builder()
    lx, ly, lz = box
    s = (1.0 / 0.0334) ** (1.0 / 3.0)          # water lattice spacing
    nx = max(1, int(round(lx / s)))
    ny = max(1, int(round(ly / s)))
    nz = max(1, int(math.ceil(nwater / (nx * ny))))
    wthick = nz * s
    z0 = 0.5 * lz - 0.5 * wthick
    sites = _lattice_sites(nx, ny, nz, (0, 0, z0), (lx, ly, z0 + wthick), rng, 0.2)
    sites = sites[rng.permutation(len(sites))][:nwater]
    wt = _water_template()
    for c in sites:
        b.add(wt, c, _random_rotation(rng))
    pt = _propanol_template()
    per_side = nprop // 2
    sp = (1.0 / 0.0075) ** (1.0 / 3.0)         # propanol lattice spacing (A)
    npx = max(1, int(round(lx / sp)))
    npy = max(1, int(round(ly / sp)))
    npz = max(1, int(math.ceil(per_side / (npx * npy))))
    pthick = npz * sp
    for side in (0, 1):
        zlo = z0 - pthick - 2.5 if side == 0 else z0 + wthick + 2.5
        ps = _lattice_sites(npx, npy, npz, (0, 0, zlo), (lx, ly, zlo + pthick), rng, 0.3)
        ps = ps[rng.permutation(len(ps))][:per_side]
        for c in ps:
            b.add_checked(pt, c, rng, box)
    return _finish(b, box, alpha=alpha, rrcut=rcut, rcut=rcut, rcut_book=rcut_book,
                   nfft=nfft, order=order, torsion_form=torsion_form, name=name, seed=seed)



def _place_layer(b, tmpl, sites, rng, box, zface, nrot=32, dmin=2.5):
    """Molecules of `tmpl` at `sites` with the orientation, out of `nrot`
    random ones, that keeps its atoms furthest (up to `dmin`) from the water
    atoms within 12 A of the plane z = zface and from the molecules already
    placed in this layer (minimum image; a KD-tree and an x-y cell hash keep
    every test local and vectorised)."""
    from scipy.spatial import cKDTree
    bx = np.asarray(box, float)
    local = tmpl["pos"] - tmpl["pos"].mean(axis=0)
    wpos = np.concatenate(b.pos, axis=0) if b.pos else np.zeros((0, 3))
    near = wpos[np.abs(wpos[:, 2] - zface) < 12.0]
    tree = cKDTree(np.mod(near, bx), boxsize=bx) if len(near) else None
    cw = 6.0
    ncx, ncy = max(1, int(bx[0] // cw)), max(1, int(bx[1] // cw))
    cells = {}
    for c in sites:
        rots = np.stack([_random_rotation(rng) for _ in range(nrot)])     # (R, 3, 3)
        cand = np.einsum("ak,rjk->raj", local, rots) + c                  # (R, A, 3)
        dm = np.full(nrot, 9.0)
        if tree is not None:
            d, _ = tree.query(np.mod(cand.reshape(-1, 3), bx), k=1)
            dm = np.minimum(dm, d.reshape(nrot, -1).min(axis=1))
        kx, ky = int(c[0] // cw) % ncx, int(c[1] // cw) % ncy
        nb = [q for dx in (-1, 0, 1) for dy in (-1, 0, 1)
              for q in cells.get(((kx + dx) % ncx, (ky + dy) % ncy), ())]
        if nb:
            q = np.concatenate(nb, axis=0)
            dq = q - c
            dq -= bx * np.round(dq / bx)
            q = q[np.all(np.abs(dq) < 9.0, axis=1)]
        if nb and len(q):
            d = cand[:, :, None, :] - q[None, None, :, :]
            d -= bx * np.round(d / bx)
            dm = np.minimum(dm, np.sqrt((d * d).sum(-1)).min(axis=(1, 2)))
        ok = np.flatnonzero(dm >= dmin)
        r = int(ok[0]) if len(ok) else int(np.argmax(dm))
        b.add(tmpl, c, rots[r])
        cells.setdefault((kx, ky), []).append(cand[r])


def interface_slab(nwater=300000, nprop=8000, box=(200.0, 200.0, 300.0), *, seed=5,
                   alpha=0.26, rcut=12.0, rcut_book=13.0, nfft=(192, 192, 288), order=6,
                   torsion_form="rb", name="interface"):
    """Two-sided alcohol/water interface: a central slab of `nwater` waters
    with `nprop` 1-propanols split over a layer on each face and vacuum beyond
    (the ~1M-atom heat-flux profiling workload at the defaults: 996,000 atoms,
    200 x 200 x 300 A, SPME 192 x 192 x 288 order 6, 12 A cutoff).  The same
    recipe at small sizes gives the downsized slab the oracle checks."""
    rng = np.random.default_rng(seed)
    b = _Builder()
    lx, ly, lz = box
    s = (1.0 / 0.0334) ** (1.0 / 3.0)          # water lattice spacing
    nx = max(1, int(round(lx / s)))
    ny = max(1, int(round(ly / s)))
    nz = max(1, int(math.ceil(nwater / (nx * ny))))
    wthick = nz * s
    z0 = 0.5 * lz - 0.5 * wthick
    sites = _lattice_sites(nx, ny, nz, (0, 0, z0), (lx, ly, z0 + wthick), rng, 0.2)
    sites = sites[rng.permutation(len(sites))][:nwater]
    wt = _water_template()
    for c in sites:
        b.add(wt, c, _random_rotation(rng))
    pt = _propanol_template()
    per_side = nprop // 2
    sp = (1.0 / 0.0075) ** (1.0 / 3.0)         # propanol lattice spacing (A)
    npx = max(1, int(round(lx / sp)))
    npy = max(1, int(round(ly / sp)))
    npz = max(1, int(math.ceil(per_side / (npx * npy))))
    pthick = npz * sp
    slab = wthick + 2.0 * (pthick + 2.5)
    if lz - slab < 5.0:
        raise ValueError(f"interface_slab: the slab ({slab:.1f} A with its alcohol layers) "
                         f"leaves {lz - slab:.1f} A of vacuum in a {lz} A box; the two "
                         "layers would meet across the periodic boundary")
    for side in (0, 1):
        zlo = z0 - pthick - 2.5 if side == 0 else z0 + wthick + 2.5
        ps = _lattice_sites(npx, npy, npz, (0, 0, zlo), (lx, ly, zlo + pthick), rng, 0.3)
        ps = ps[rng.permutation(len(ps))][:per_side]
        _place_layer(b, pt, ps, rng, box, z0 if side == 0 else z0 + wthick)
    return _finish(b, box, alpha=alpha, rrcut=rcut, rcut=rcut, rcut_book=rcut_book,
                   nfft=nfft, order=order, torsion_form=torsion_form, name=name, seed=seed)

def mixture_box(n_target, composition, *, seed=4, density_gcc=0.997, rcut=10.0,
                rcut_book=11.0, order=4, alpha=0.32, spacing=1.0, torsion_form="rb",
                alcohol_frac=0.1, name="mix"):
    """Cubic replica of about `n_target` atoms: 'water', 'butane' or 'mix'
    (water + propanol at `alcohol_frac` mole fraction), at the given density,
    PME grid spacing <= `spacing` rounded up to 2^a 3^b 5^c."""
    rng = np.random.default_rng(seed)
    tm = {"water": [("water", 1.0)], "butane": [("butane", 1.0)],
          "mix": [("water", 1.0 - alcohol_frac), ("propanol", alcohol_frac)]}[composition]
    tmpls = {k: TEMPLATES[k]() for k, _ in tm}
    mean_atoms = sum(f * len(tmpls[k]["types"]) for k, f in tm)
    nmol = max(1, int(round(n_target / mean_atoms)))
    kinds = []
    for k, f in tm:
        kinds += [k] * int(round(nmol * f))
    kinds = kinds[:nmol] + [tm[0][0]] * max(0, nmol - len(kinds))
    kinds = [kinds[i] for i in rng.permutation(len(kinds))]
    mass = sum(sum(ATOM_TYPES[TYPE_INDEX[t] - 1].mass for t in tmpls[k]["types"]) for k in kinds)
    if composition == "butane":
        density_gcc = min(density_gcc, 0.60)         # liquid n-butane
    vol = mass / 6.0221367e23 / density_gcc * 1e24   # A^3
    box = vol ** (1.0 / 3.0)
    box = max(box, 2.0 * rcut_book + 0.5)
    nbig = sum(1 for k in kinds if k != "water")
    nwat = nmol - nbig
    nsite = int(math.ceil(max(nbig, 1) ** (1.0 / 3.0)))
    sites = _lattice_sites(nsite, nsite, nsite, (0, 0, 0), (box, box, box), rng, 0.2)
    sites = sites[rng.permutation(len(sites))][:nbig]
    b = _Builder()
    bx = np.array([box, box, box])
    # larger molecules first (orientation search against what is placed), then
    # waters by random sequential addition on a fine jittered grid, so the
    # replica keeps its drawn composition and size
    big = [k for k in kinds if k != "water"]
    for k, c in zip(big, sites):
        b.add_checked(tmpls[k], c, rng, bx, dmin=2.5, tries=400)
    heavy = np.concatenate(b.pos, axis=0) if b.pos else np.zeros((0, 3))
    if nbig == 0:   # pure water: a jittered lattice holds every molecule
        m = int(math.ceil(nwat ** (1.0 / 3.0)))
        wsites = _lattice_sites(m, m, m, (0, 0, 0), (box, box, box), rng, 0.2)
        wsites = wsites[rng.permutation(len(wsites))][:nwat]
    else:
        wsites = _fill_sites(nwat, box, heavy, rng)
    for c in wsites:
        b.add(tmpls["water"], c, _random_rotation(rng))
    g = fft_friendly(int(math.ceil(box / spacing)))
    return _finish(b, (box, box, box), alpha=alpha, rrcut=rcut, rcut=rcut,
                   rcut_book=rcut_book, nfft=(g, g, g), order=order,
                   torsion_form=torsion_form, name=name, seed=seed)


def _fill_sites(n, box, solute, rng, d_solute=2.1, d_water=2.3, grid=0.9):
    """Up to `n` centres in the periodic cube [0, box)^3 picked in random order
    from a jittered grid of spacing ~`grid`, each at least `d_solute` from every
    solute atom and `d_water` from every centre already taken (a spatial hash
    of cell size >= both distances keeps each test local)."""
    m = max(1, int(box / grid))
    cand = _lattice_sites(m, m, m, (0, 0, 0), (box, box, box), rng, 0.3 * box / m)
    cand = np.mod(cand[rng.permutation(len(cand))], box)
    nc = max(1, int(box / max(d_solute, d_water)))
    cw = box / nc
    cells = {}

    def key(p):
        return tuple((np.floor(p / cw).astype(int)) % nc)

    def near(p, dmin, kind):
        k = key(p)
        for dx in (-1, 0, 1):
            for dy in (-1, 0, 1):
                for dz in (-1, 0, 1):
                    for q, t in cells.get(((k[0] + dx) % nc, (k[1] + dy) % nc,
                                           (k[2] + dz) % nc), ()):
                        d = q - p
                        d -= box * np.round(d / box)
                        lim = d_solute if t else d_water
                        if d @ d < lim * lim:
                            return True
        return False

    for q in solute:
        q = np.mod(q, box)
        cells.setdefault(key(q), []).append((q, True))
    out = []
    for p in cand:
        if len(out) == n:
            break
        if near(p, d_water, False):
            continue
        cells.setdefault(key(p), []).append((p, False))
        out.append(p)
    return out


def fft_friendly(n: int) -> int:
    """Smallest m >= n with m = 2^a 3^b 5^c."""
    m = max(1, n)
    while True:
        k = m
        for p in (2, 3, 5):
            while k % p == 0:
                k //= p
        if k == 1:
            return m
        m += 1


def _replica(args):
    n, kind, rseed, frac, name = args
    return mixture_box(n, kind, seed=rseed, alcohol_frac=frac, name=name)


def replica_ensemble(nrep=64, nmin=1500, nmax=12000, seed=4, workers=1, mp_context="fork"):
    """Uneven Green-Kubo replicas: sizes log-uniform in [nmin, nmax], each a
    water, n-butane or water/propanol (5-20 mol%) box with its own PME grid.
    The replicas' parameters are drawn in order from one generator, so
    `workers` > 1 (forked processes building replicas side by side) gives the
    same replicas.  Use `mp_context` "fork" only before anything touches the
    GPU; "spawn" starts fresh interpreters and is safe afterwards."""
    rng = np.random.default_rng(seed)
    jobs = []
    for r in range(nrep):
        n = int(round(math.exp(rng.uniform(math.log(nmin), math.log(nmax)))))
        kind = ["water", "butane", "mix"][int(rng.integers(0, 3))]
        frac = float(rng.uniform(0.05, 0.20))
        jobs.append((n, kind, int(rng.integers(1 << 30)), frac, f"rep{r}-{kind}"))
    if workers <= 1:
        return [_replica(j) for j in jobs]
    import multiprocessing as mp
    # largest first, so the pool's tail is short
    order = sorted(range(nrep), key=lambda r: -jobs[r][0])
    with mp.get_context(mp_context).Pool(workers) as pool:
        built = pool.map(_replica, [jobs[r] for r in order], chunksize=1)
    reps = [None] * nrep
    for r, sysd in zip(order, built):
        reps[r] = sysd
    return reps


def exclusions(sysd):
    """Exclusion CSR in the reference's final mkexcl layout (mkexcl.F90:193-222):
    1-based, per atom the sorted unique partners j > i from bond (1-2),
    angle (1-3) and torsion (1-4) end atoms; an atom with none holds a single 0."""
    n = sysd["natom"]
    part = [set() for _ in range(n + 1)]
    for i, j in sysd["bond_ij"]:
        part[i].add(j)
        part[j].add(i)
    for i, _j, k in sysd["angl_ijk"]:
        part[i].add(k)
        part[k].add(i)
    for key in ("tors_ijkl", "torsrb_ijkl"):
        for i, _j, _k, l in sysd[key]:
            part[i].add(l)
            part[l].add(i)
    idx = np.zeros(n + 1, np.int32)
    lst = []
    for a in range(1, n + 1):
        idx[a - 1] = len(lst) + 1
        js = sorted(x for x in part[a] if x > a)
        lst += js if js else [0]
    idx[n] = len(lst) + 1
    return idx, np.array(lst, np.int32)
