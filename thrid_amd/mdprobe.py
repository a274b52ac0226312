"""Time device-resident MD steps of the standard workloads, with and without
the heat-flux sample, one context at a time (the workload a kernel profile
of the heat-flux path is taken on).

    python -m thrid_amd.mdprobe replicas   # three 1.6k-11k-atom replicas, bulk flux
    python -m thrid_amd.mdprobe slab       # the 18.4k water/propanol slab, bulk and 1 A bins
    python -m thrid_amd.mdprobe interface  # the ~1M-atom interface, 300 bins, both thermostats

Each line printed: workload, precision, atoms, flux form, microseconds per
MD step (wall time around md_run after a short warm-up).
"""
import sys
import time

import numpy as np

from thrid_amd import abi, md, systems
from thrid_amd.pgk import ForceContext


def _relaxed(s, **kw):
    cons = md.rigid_constraints(s)
    c = ForceContext(s, "fp64")
    s["pos"] = md.relax(c, s, cons, **kw)
    c.close()
    return cons


def _time(c, s, nstep, nbook):
    c.md_run(3, nbook, s["rcut_book"], 3)
    c.synchronize()
    t0 = time.perf_counter()
    c.md_run(nstep, nbook, s["rcut_book"], nstep)
    c.synchronize()
    return (time.perf_counter() - t0) / nstep * 1e6


def _mole_waters(s):
    ht = np.ones(int(s["natom"]), np.int32)
    ht[np.isin(s["atype"], [systems.TYPE_INDEX["OB"], systems.TYPE_INDEX["HB"]])] = abi.HFTYP_MOLE
    return ht


def replicas(prec="mixed", nstep=50):
    """Three replicas spanning the 64-replica ensemble's 1.5k-12k-atom range
    (water 1.6k, water/propanol 5k, n-butane 11k atoms): no flux, then the
    bulk flux every step."""
    reps = [systems.mixture_box(1600, "water", seed=21, name="water1600"),
            systems.mixture_box(5000, "mix", seed=22, alcohol_frac=0.12, name="mix5000"),
            systems.mixture_box(11000, "butane", seed=23, name="butane11000")]
    for s in reps:
        cons = _relaxed(s)
        c = ForceContext(s, prec)
        c.md_setup(s["mass"], *cons, 1.0)
        c.md_set_velocities(md.maxwell_boltzmann(s, 300.0, seed=5))
        print(f"{s['name']} {prec} {s['natom']} none {_time(c, s, nstep, 10):.1f}", flush=True)
        c.md_heatf(dict(mode=abi.HF_BULK, z1=np.array([0.0]), z2=np.array([s["box"][2]])), 1)
        print(f"{s['name']} {prec} {s['natom']} bulk {_time(c, s, nstep, 10):.1f}", flush=True)
        c.close()


def slab(prec="mixed", nstep=20):
    """The 18.4k-atom water/propanol slab (50 x 50 x 150 A): no flux, the bulk
    flux, and 1 A bins over the box (waters by molecule centre), every step."""
    s = systems.propanol_water_slab()
    cons = _relaxed(s)
    lz = float(s["box"][2])
    ht = _mole_waters(s)
    z = np.arange(0.0, lz + 0.5, 1.0)
    c = ForceContext(s, prec)
    c.md_setup(s["mass"], *cons, 1.0)
    c.md_set_velocities(md.maxwell_boltzmann(s, 300.0, seed=5))
    n = s["natom"]
    print(f"slab {prec} {n} none {_time(c, s, nstep, 20):.1f}", flush=True)
    c.md_heatf(dict(mode=abi.HF_BULK, z1=np.array([0.0]), z2=np.array([lz]), hftyp=ht), 1)
    print(f"slab {prec} {n} bulk {_time(c, s, nstep, 20):.1f}", flush=True)
    c.md_heatf(dict(mode=abi.HF_VOLUME, z1=z[:-1], z2=z[1:], hftyp=ht), 1)
    print(f"slab {prec} {n} bins{len(z) - 1} {_time(c, s, nstep, 20):.1f}", flush=True)
    c.close()


def interface(prec="mixed", nstep=5):
    """The ~1M-atom two-sided alcohol/water interface (200 x 200 x 300 A) with
    Langevin ends, the chain on the rest and the 300-bin flux every step."""
    s = systems.interface_slab()
    cons = _relaxed(s, schedule=((0.02, 10), (0.1, 10), (0.5, 10)))
    n = s["natom"]
    lx, ly, lz = (float(b) for b in s["box"])
    ij, d, gs = cons
    z = np.arange(0.0, lz + 0.5, 1.0)
    zmin, zmax = float(s["pos"][:, 2].min()), float(s["pos"][:, 2].max())
    ends = np.array([[0.0, lx, 0.0, ly, zmin - 1.0, zmin + 20.0],
                     [0.0, lx, 0.0, ly, zmax - 20.0, zmax + 1.0]])
    c = ForceContext(s, prec)
    c.md_setup(s["mass"], ij, d, gs, 1.0)
    c.md_set_velocities(md.maxwell_boltzmann(s, 300.0, seed=55))
    c.md_langevin(ends, [320.0, 280.0], [0.01, 0.01], seed=5)
    c.md_nhc(300.0, 0.01, 3 * n - len(d) - 3)
    c.md_heatf(dict(mode=abi.HF_VOLUME, z1=z[:-1], z2=z[1:], hftyp=_mole_waters(s)), 1)
    print(f"interface {prec} {n} bins{len(z) - 1} {_time(c, s, nstep, 20):.1f}", flush=True)
    c.close()


def main(argv):
    cases = {"replicas": replicas, "slab": slab, "interface": interface}
    if not argv or argv[0] not in cases:
        print(__doc__)
        return 2
    prec = argv[1] if len(argv) > 1 else "mixed"
    cases[argv[0]](prec)
    return 0


if __name__ == "__main__":
    sys.exit(main(sys.argv[1:]))
