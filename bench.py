"""Benchmark: MD steps of the 100,001-atom SPME water box on one MI355X.

Workload: 33,333 rigid 3-site waters plus one Na+/Cl- pair in a 100.1 A
cubic box, SPME 128^3 order 4, alpha 0.32/A, 10 A real-space and LJ cutoff,
rcut_book 11 A, dt = 1 fs.  One step = one device-resident NVE step
(velocity Verlet, every force term on MTS_LONG, RATTLE on the waters); the
Verlet list is rebuilt every NSTEP_BOOK steps inside the timed region.  The
start is the generated lattice relaxed by a short untimed FP64 run with
velocity rescaling to 300 K.

Prints one JSON line.  value = atom-steps/s of the mixed form; beside it the
FP64 form, ns/day, the force-only step rate, the pair kernel's roofline (HIP
events on the library stream), each form's force error against the
reference's FP64 forces, NVE energy drift per form and the reference CPU
baseline.  Further legs (each with its own flag): the heat-flux kernels on
the water box, the 18.4k-atom water/propanol slab and the ~1M-atom interface
under their thermostats with the heat flux every step, and the replica
ensemble.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

NSTEP_BOOK = 20
DT_FS = 1.0
# energies and virials are sampled every OUT_INTERVAL steps of the timed run,
# the output interval of the reference's own peachgk.ini (outinterval and
# pressinterval 100); steps between samples evaluate forces only
OUT_INTERVAL = 100
PEAK = {"fp64": 78.6, "mixed": 157.3, "fp32": 157.3}   # TFLOP/s, MI355X vector


def pair_counts(sysd):
    """In-cutoff half pairs (Coulomb, r <= rrcut) and LJ pairs with eps != 0
    (r <= rcut) of the listed set (exclusions removed) at the current positions."""
    from scipy.spatial import cKDTree
    from thrid_amd.systems import exclusions
    box = sysd["box"]
    w = sysd["pos"] - box * np.floor(sysd["pos"] / box)
    w = np.minimum(w, box * (1 - 1e-12))
    t = cKDTree(w, boxsize=box)
    rc = max(sysd["rrcut"], sysd["rcut"])
    pr = t.query_pairs(rc, output_type="ndarray")
    d = w[pr[:, 0]] - w[pr[:, 1]]
    d -= box * np.round(d / box)
    r = np.sqrt((d * d).sum(1))
    idx, lst = exclusions(sysd)
    n = sysd["natom"]
    owner = np.repeat(np.arange(n, dtype=np.int64), np.diff(idx).astype(np.int64))
    sel = lst > 0
    exk = owner[sel] * n + (lst[sel].astype(np.int64) - 1)
    a = np.minimum(pr[:, 0], pr[:, 1]).astype(np.int64)
    b = np.maximum(pr[:, 0], pr[:, 1]).astype(np.int64)
    keys = a * n + b
    keep = ~np.isin(keys, exk)
    ncoul = int(np.count_nonzero(keep & (r <= sysd["rrcut"])))
    ti = sysd["atype"][a] - 1
    tj = sysd["atype"][b] - 1
    eps = sysd["lj_eps"][ti, tj]
    nlj = int(np.count_nonzero(keep & (r <= sysd["rcut"]) & (eps != 0.0)))
    return ncoul, nlj


def equilibrated_state(sysd, seed=5):
    """Relax the generated start: 2000 FP64 steps with velocity rescaling to
    300 K (untimed).  Returns positions, velocities and the constraints."""
    from thrid_amd import md
    from thrid_amd.pgk import ForceContext
    ij, d, gs = md.rigid_constraints(sysd)
    ctx = ForceContext(sysd, "fp64")
    ctx.md_setup(sysd["mass"], ij, d, gs, DT_FS)
    ctx.md_set_velocities(md.maxwell_boltzmann(sysd, 300.0, seed=seed))
    pos, vel = md.equilibrate(ctx, sysd, 300.0, 10, 200, NSTEP_BOOK, len(d))
    ctx.close()
    return pos, vel, (ij, d, gs)


def run_md(sysd, state, precision, steps, warmup, ndrift):
    """Time exactly `steps` MD steps (after `warmup`), then an NVE run of
    `ndrift` steps for the energy drift.  Returns (s/step, stage ms, list
    info, drift eref/ns/atom, energy rows of the drift run)."""
    from thrid_amd import md
    from thrid_amd.pgk import ForceContext
    pos, vel, (ij, d, gs) = state
    s2 = dict(sysd)
    s2["pos"] = pos
    ctx = ForceContext(s2, precision)
    ctx.md_setup(sysd["mass"], ij, d, gs, DT_FS)
    ctx.md_set_velocities(vel)
    if warmup:
        ctx.md_run(warmup, NSTEP_BOOK, sysd["rcut_book"], warmup)
    ctx.synchronize()
    ctx.profile_stages(1)          # HIP events around the pair kernel only
    ctx.profile(1)
    nsample = OUT_INTERVAL if steps % OUT_INTERVAL == 0 else steps
    t0 = time.perf_counter()
    ctx.md_run(steps, NSTEP_BOOK, sysd["rcut_book"], nsample)
    ctx.synchronize()
    t1 = time.perf_counter()
    ms, calls = ctx.profile(0)
    info = ctx.list_info()
    # stage breakdown from a separate profiled run of the same length
    ctx.profile_stages(0xff)
    ctx.profile(1)
    ctx.md_run(steps, NSTEP_BOOK, sysd["rcut_book"], nsample)
    ms_all, calls_all = ctx.profile(0)
    ctx.set_positions(pos)
    ctx.md_set_velocities(vel)
    rows = ctx.md_run(ndrift, NSTEP_BOOK, sysd["rcut_book"], 50)
    ctx.close()
    stages = ["pairs", "excl", "spread", "fft", "gather", "bonded", "list", "other"]
    prof = {s: (float(ms_all[k] / calls_all[k]) if calls_all[k] else 0.0)
            for k, s in enumerate(stages)}
    prof["pairs_timed_region"] = float(ms[0] / calls[0]) if calls[0] else 0.0
    return (t1 - t0) / steps, prof, info, md.drift_per_ns(rows, 50, DT_FS, sysd["natom"]), rows


def run_force_only(sysd, precision, steps):
    """Force-only steps at the generated positions (queued without host
    round trips) and the forces there, for the accuracy check."""
    from thrid_amd.pgk import ForceContext
    ctx = ForceContext(sysd, precision)
    ctx.mklist(sysd["rcut_book"])
    for _ in range(3):
        ctx.calforce_async()
    ctx.synchronize()
    t0 = time.perf_counter()
    for k in range(steps):
        if k % NSTEP_BOOK == 0:
            ctx.mklist(sysd["rcut_book"])
        ctx.calforce_async()
    ctx.energies()
    t1 = time.perf_counter()
    force = ctx.calforce()["force"]
    ctx.close()
    return (t1 - t0) / steps, force


def run_heatf(sysd, state, precision, reps=5):
    """A heat-flux step of the 100k box: calforce_hf with the per-atom
    arrays left in HBM, then calheatf on the device, for one bulk region
    (the Green-Kubo case) and for 10 volume slabs; beside it the same
    calforce_hf copying the per-atom arrays to the host, which a host-side
    calheatf needs, and the fused calforce_heatf.  Wall-clock ms per call
    (each call waits for its results)."""
    from thrid_amd import abi, md
    from thrid_amd.pgk import ForceContext
    pos, vel, _cons = state
    s2 = dict(sysd)
    s2["pos"] = pos
    ctx = ForceContext(s2, precision)
    ctx.mklist(sysd["rcut_book"])
    mass = np.asarray(sysd["mass"], np.float64) / md.MD_CONV    # m v^2 / 2 in eref
    lz = sysd["box"][2]
    edges = np.linspace(0.0, lz, 11)
    cases = {"bulk_1": dict(mode=abi.HF_BULK, z1=np.array([0.0]), z2=np.array([lz])),
             "volume_10": dict(mode=abi.HF_VOLUME, z1=edges[:-1], z2=edges[1:])}
    out = {}
    for name, hf in cases.items():
        ctx.calforce_hf(hf, per_atom=False)
        ctx.calheatf(vel, mass)
        t0 = time.perf_counter()
        for _ in range(reps):
            ctx.calforce_hf(hf, per_atom=False)
        t1 = time.perf_counter()
        for _ in range(reps):
            ctx.calheatf(vel)
        t2 = time.perf_counter()
        for _ in range(reps):
            ctx.calforce_hf(hf, per_atom=True)
        t3 = time.perf_counter()
        ctx.calforce_heatf(hf, vel)
        t4 = time.perf_counter()
        for _ in range(reps):
            ctx.calforce_heatf(hf, vel)
        t5 = time.perf_counter()
        out[name] = {"calforce_hf_ms": (t1 - t0) / reps * 1e3,
                     "calheatf_ms": (t2 - t1) / reps * 1e3,
                     "calforce_hf_with_host_arrays_ms": (t3 - t2) / reps * 1e3,
                     "fused_calforce_heatf_ms": (t5 - t4) / reps * 1e3}
    ctx.close()
    return out


def run_extras(sysd, state, precision, steps=200):
    """Cost of the region Langevin thermostat in the MD step (two slabs
    holding a fifth of the atoms; ms per step with and without it) and of
    one per-atom SPME reciprocal virial call (six transform pairs, with the
    copy of the 9N doubles to the host)."""
    from thrid_amd import abi
    from thrid_amd.pgk import ForceContext
    pos, vel, (ij, d, gs) = state
    s2 = dict(sysd)
    s2["pos"] = pos
    lx, ly, lz = (float(b) for b in sysd["box"])
    regions = [[0, lx, 0, ly, 0, 0.1 * lz], [0, lx, 0, ly, 0.5 * lz, 0.6 * lz]]
    out = {}
    for lang in (False, True):
        ctx = ForceContext(s2, precision)
        ctx.md_setup(sysd["mass"], ij, d, gs, DT_FS)
        ctx.md_set_velocities(vel)
        if lang:
            ctx.md_langevin(regions, [350.0, 250.0], [0.01, 0.01], seed=1)
        ctx.md_run(20, NSTEP_BOOK, sysd["rcut_book"], 20)
        ctx.synchronize()
        t0 = time.perf_counter()
        ctx.md_run(steps, NSTEP_BOOK, sysd["rcut_book"], steps)
        ctx.synchronize()
        out["langevin_ms_per_step" if lang else "nve_ms_per_step"] = \
            (time.perf_counter() - t0) / steps * 1e3
        ctx.close()
    ctx = ForceContext(s2, precision)
    ctx.mklist(sysd["rcut_book"])
    ctx.calforce(abi.MTS_LONG)
    ctx.recip_virial_atom()
    t0 = time.perf_counter()
    for _ in range(5):
        ctx.recip_virial_atom()
    out["recip_virial_atom_ms"] = (time.perf_counter() - t0) / 5 * 1e3
    ctx.close()
    return out


def run_interface(steps, precs):
    """The ~1M-atom two-sided alcohol/water interface (300,000 waters + 8,000
    1-propanols, 200 x 200 x 300 A, SPME 192 x 192 x 288 order 6, 12 A cutoff):
    after a short FP64 quench of the generated start, `steps` MD steps per form
    with Langevin thermostats (320 K / 280 K) on the 20 A at each end of the
    slab, a Nose-Hoover chain (300 K) on the rest, and the 1 A binned heat flux
    (300 volume regions, waters by molecule centre) sampled every step.
    Returns ms per step and atom-steps/s per form."""
    from thrid_amd import abi, md, systems
    from thrid_amd.pgk import ForceContext
    t0 = time.perf_counter()
    s = systems.interface_slab()
    cons = md.rigid_constraints(s)
    c = ForceContext(s, "fp64")
    s["pos"] = md.relax(c, s, cons, schedule=((0.02, 10), (0.1, 10), (0.5, 10)))
    c.close()
    t_setup = time.perf_counter() - t0
    n = s["natom"]
    lx, ly, lz = (float(b) for b in s["box"])
    ij, d, gs = cons
    ht = np.ones(n, np.int32)
    ht[np.isin(s["atype"], [systems.TYPE_INDEX["OB"], systems.TYPE_INDEX["HB"]])] = abi.HFTYP_MOLE
    z = np.arange(0.0, lz + 0.5, 1.0)
    zmin, zmax = float(s["pos"][:, 2].min()), float(s["pos"][:, 2].max())
    ends = np.array([[0.0, lx, 0.0, ly, zmin - 1.0, zmin + 20.0],
                     [0.0, lx, 0.0, ly, zmax - 20.0, zmax + 1.0]])
    out = {"atoms": n, "setup_s": t_setup, "steps": steps, "heat_flux_bins": len(z) - 1}
    for prec in precs:
        c = ForceContext(s, prec)
        c.md_setup(s["mass"], ij, d, gs, DT_FS)
        c.md_set_velocities(md.maxwell_boltzmann(s, 300.0, seed=55))
        c.md_langevin(ends, [320.0, 280.0], [0.01, 0.01], seed=5)
        c.md_nhc(300.0, 0.01, 3 * n - len(d) - 3)
        c.md_heatf(dict(mode=abi.HF_VOLUME, z1=z[:-1], z2=z[1:], hftyp=ht), 1)
        c.md_run(2, NSTEP_BOOK, s["rcut_book"], 2)
        c.synchronize()
        t0 = time.perf_counter()
        c.md_run(steps, NSTEP_BOOK, s["rcut_book"], steps)
        c.synchronize()
        sec = (time.perf_counter() - t0) / steps
        rows = c.md_heatf_series()
        out[prec] = {"ms_per_step": sec * 1e3, "atom_steps_per_s": n / sec,
                     "ns_per_day": 86400.0 / sec * DT_FS * 1e-6,
                     "flux_rows": int(rows.shape[0]),
                     "flux_finite": bool(np.all(np.isfinite(rows)))}
        c.close()
    return out


def run_slab(steps, precs):
    """The 18.4k-atom water/propanol slab (4096 waters, 512 1-propanols,
    50 x 50 x 150 A, SPME 64 x 64 x 192 order 4, 12 A cutoff): after a short
    FP64 quench, `steps` MD steps per form under a Nose-Hoover chain at 300 K
    with the box's heat-flux vector (bulk form, one region over the box, atom
    centres) sampled every step.  Returns ms per step and atom-steps/s per form."""
    from thrid_amd import abi, md, systems
    from thrid_amd.pgk import ForceContext
    t0 = time.perf_counter()
    s = systems.propanol_water_slab()
    cons = md.rigid_constraints(s)
    c = ForceContext(s, "fp64")
    s["pos"] = md.relax(c, s, cons)
    c.close()
    n = s["natom"]
    ij, d, gs = cons
    lz = float(s["box"][2])
    out = {"atoms": n, "setup_s": time.perf_counter() - t0, "steps": steps}
    for prec in precs:
        c = ForceContext(s, prec)
        c.md_setup(s["mass"], ij, d, gs, DT_FS)
        c.md_set_velocities(md.maxwell_boltzmann(s, 300.0, seed=57))
        c.md_nhc(300.0, 0.01, 3 * n - len(d) - 3)
        c.md_heatf(dict(mode=abi.HF_BULK, z1=np.array([0.0]), z2=np.array([lz])), 1)
        c.md_run(5, NSTEP_BOOK, s["rcut_book"], 5)
        c.synchronize()
        t0 = time.perf_counter()
        c.md_run(steps, NSTEP_BOOK, s["rcut_book"], steps)
        c.synchronize()
        sec = (time.perf_counter() - t0) / steps
        rows = c.md_heatf_series()
        out[prec] = {"ms_per_step": sec * 1e3, "atom_steps_per_s": n / sec,
                     "ns_per_day": 86400.0 / sec * DT_FS * 1e-6,
                     "flux_rows": int(rows.shape[0]),
                     "flux_finite": bool(np.all(np.isfinite(rows)))}
        c.close()
    return out


def copy_calibration(nbytes=1 << 30, reps=5):
    """A plain device copy of known size (scale for the HBM byte counters)."""
    import torch
    src = torch.empty(nbytes // 4, dtype=torch.float32, device="cuda").fill_(1.0)
    dst = torch.empty_like(src)
    for _ in range(reps):
        dst.copy_(src)
    torch.cuda.synchronize()
    print(f"# copy calibration: {reps} copies of {nbytes} bytes", file=sys.stderr, flush=True)


def traffic_from_profiles(kernel_prefix):
    """HBM bytes per launch of the kernel from the committed PMC summary
    (FETCH_SIZE + WRITE_SIZE, each scaled by the copy-kernel factor)."""
    p = os.path.join(ROOT, "profiles", "pmc_traffic.json")
    try:
        with open(p) as f:
            d = json.load(f)
        return d["kernels"][kernel_prefix]["hbm_bytes_per_launch"]
    except (OSError, KeyError, ValueError):
        return None


def cpu_baseline(sysd, budget_s=20.0):
    """The reference's own calforce (compiled from source, oracle/_ref) on one
    core, timed on a bounded number of calls; falls back to the C restatement.
    Also returns its forces at the benchmark positions (the FP64 checker)."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import pyoracle
    from thrid_amd import abi
    kind = "reference" if pyoracle.reference_available() else "port"
    m = pyoracle.Reference(sysd) if kind == "reference" else pyoracle.Oracle(sysd)
    timed = kind == "reference" and os.path.abspath(pyoracle.REF_SO) == \
        os.path.abspath(pyoracle.REF_TIMED_SO)
    if timed:
        m.times(reset=True)
    t0 = time.perf_counter()
    m.mklist(sysd["rcut_book"], 1)
    t_list = time.perf_counter() - t0
    ncall, t_force, force = 0, 0.0, None
    while t_force < budget_s and ncall < 50:
        t0 = time.perf_counter()
        r = m.calforce(abi.MTS_LONG)
        t_force += time.perf_counter() - t0
        ncall += 1
        force = r["force"]
    per_step = t_force / ncall + t_list / NSTEP_BOOK
    base = dict(value=sysd["natom"] / per_step, unit="atom-steps/s", cores=1, kind=kind,
                cpu=cpu_model(),
                sample=f"{ncall} calforce calls + 1 cell-list build (amortised over "
                       f"{NSTEP_BOOK} steps) of the same 100,001-atom box, serial, "
                       f"list {t_list:.2f} s, calforce {t_force / ncall:.2f} s/call")
    if timed:
        # the reference's own -DTIME_M split (time_method.F90:157-265), CPU
        # seconds per calforce call (drct, rcpr, bond) and per list build (table)
        t = m.times()
        base["time_m_split_s"] = {k: (t[k] / ncall if k != "table_time" else t[k])
                                  for k in ("drct_time", "rcpr_time", "bond_time", "table_time")}
    return base, force


def cpu_model():
    """The host CPU's model name and the logical CPUs this process may use."""
    name = "unknown"
    try:
        with open("/proc/cpuinfo") as f:
            for ln in f:
                if ln.startswith("model name"):
                    name = ln.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    return f"{name} ({len(os.sched_getaffinity(0))} logical CPUs available, 1 used)"


def cpu_leg_subprocess():
    """Run cpu_baseline in a child process: the compiled reference writes its
    own progress lines to stdout, which must not mix with the JSON line."""
    import subprocess
    import tempfile
    with tempfile.TemporaryDirectory() as d:
        log = os.path.join(d, "cpu.log")
        env = dict(os.environ)
        timed = os.path.join(ROOT, "oracle", "_ref", "libpgkref_timed.so")
        if os.path.exists(timed):   # the reference with its -DTIME_M timers
            env["PGK_REF_SO"] = timed
        with open(log, "w") as lf:
            r = subprocess.run([sys.executable, os.path.abspath(__file__), "--cpu-leg", d],
                               stdout=lf, stderr=subprocess.STDOUT, env=env)
        if r.returncode != 0:
            sys.stderr.write(open(log).read()[-4000:])
            raise RuntimeError(f"CPU baseline leg failed ({r.returncode})")
        with open(os.path.join(d, "cpu.json")) as f:
            base = json.load(f)
        fref = np.load(os.path.join(d, "force.npy"))
    return base, fref


def force_error(f, ref):
    d = np.sqrt(((f - ref) ** 2).sum(1))
    rms_ref = np.sqrt(np.mean((ref ** 2).sum(1)))
    return {"rms_rel": float(np.sqrt(np.mean(d ** 2)) / rms_ref),
            "max_rel": float(d.max() / rms_ref)}


def run_ensemble(reps, steps, warmup, hf_every):
    """The batched Green-Kubo ensemble leg: every replica relaxed from its
    generated start, then `steps` mixed-precision MD steps of all replicas run
    one replica after another and then stepped together (pgk_md_run_batch),
    with the box's heat-flux vector (bulk form, one region) sampled every
    `hf_every` steps on the device (0: none).  Returns the two throughputs in
    atom-steps/s."""
    from thrid_amd import abi, md
    from thrid_amd.pgk import ForceContext, md_run_batch
    ctxs, cons = [], []
    t0 = time.perf_counter()
    for k, s in enumerate(reps):
        cn = md.rigid_constraints(s)
        c = ForceContext(s, "mixed")
        md.relax(c, s, cn)
        c.md_setup(s["mass"], *cn, DT_FS)
        c.md_set_velocities(md.maxwell_boltzmann(s, 300.0, seed=100 + k))
        if hf_every > 0:
            c.md_heatf(dict(mode=abi.HF_BULK, z1=np.array([0.0]), z2=np.array([s["box"][2]])),
                       hf_every)
        ctxs.append(c)
        cons.append(cn)
    t_setup = time.perf_counter() - t0
    books = [s["rcut_book"] for s in reps]
    natoms = sum(s["natom"] for s in reps)
    for c, s in zip(ctxs, reps):
        c.md_run(warmup, NSTEP_BOOK, s["rcut_book"], warmup)
    for c in ctxs:
        c.synchronize()
    t0 = time.perf_counter()
    for c, s in zip(ctxs, reps):
        c.md_run(steps, NSTEP_BOOK, s["rcut_book"], steps)
    for c in ctxs:
        c.synchronize()
    t_seq = time.perf_counter() - t0
    md_run_batch(ctxs, warmup, NSTEP_BOOK, books, warmup)
    t0 = time.perf_counter()
    md_run_batch(ctxs, steps, NSTEP_BOOK, books, steps)
    t_batch = time.perf_counter() - t0
    nhf = [c.md_heatf_series().shape[0] for c in ctxs] if hf_every > 0 else [0]
    for c in ctxs:
        c.close()
    return {"replicas": len(reps), "atoms": natoms,
            "sizes": [int(min(s["natom"] for s in reps)), int(max(s["natom"] for s in reps))],
            "steps": steps, "precision": "mixed",
            "heat_flux_every": hf_every,
            "heat_flux_rows_per_replica": int(min(nhf)),
            "sequential_atom_steps_per_s": natoms * steps / t_seq,
            "batched_atom_steps_per_s": natoms * steps / t_batch,
            "batched_over_sequential": t_seq / t_batch,
            "setup_s": t_setup,
            "method": "one context and stream per replica; step s of every replica queued "
                      "before step s+1 of any"}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--precision", choices=["all", "both", "mixed", "fp64"], default="all",
                    help="all: fp64, mixed and fp32 (the three rungs of the precision "
                         "ladder); both: fp64 and mixed")
    ap.add_argument("--drift-steps", type=int, default=3000,
                    help="NVE steps per form for the energy drift")
    ap.add_argument("--copy-calib", action="store_true",
                    help="first run a plain 1 GiB device copy (counter calibration)")
    ap.add_argument("--ensemble", type=int, default=64, metavar="NREP",
                    help="replicas in the batched-ensemble leg (0 skips it)")
    ap.add_argument("--ensemble-steps", type=int, default=100)
    ap.add_argument("--ensemble-hf", type=int, default=1, metavar="K",
                    help="heat-flux sample every K ensemble steps (0: none)")
    ap.add_argument("--no-heatf", action="store_true", help="skip the heat-flux leg")
    ap.add_argument("--slab", type=int, default=200, metavar="STEPS",
                    help="MD steps of the 18.4k-atom slab leg per form (0 skips it)")
    ap.add_argument("--interface", type=int, default=20, metavar="STEPS",
                    help="MD steps of the ~1M-atom interface leg per form (0 skips it)")
    ap.add_argument("--extras", action="store_true",
                    help="add the region Langevin and per-atom virial cost leg")
    ap.add_argument("--cpu-leg", metavar="DIR", help=argparse.SUPPRESS)
    args = ap.parse_args()
    if args.cpu_leg:   # child process of the CPU baseline leg (no GPU use)
        from thrid_amd import systems
        base, fref = cpu_baseline(systems.water_ion_box_100k())
        with open(os.path.join(args.cpu_leg, "cpu.json"), "w") as f:
            json.dump(base, f)
        np.save(os.path.join(args.cpu_leg, "force.npy"), fref)
        return
    if args.gpus != 1:
        print("bench.py measures one GPU: this build runs the force path on a single "
              "MI355X (no multi-GPU path); use --gpus 1", file=sys.stderr)
        sys.exit(2)
    from thrid_amd import systems
    reps = None
    if args.ensemble > 0:
        # built by forked workers, before anything in this process uses the GPU
        reps = systems.replica_ensemble(nrep=args.ensemble,
                                        workers=min(16, os.cpu_count() or 1))
    if args.copy_calib:
        copy_calibration()
    sysd = systems.water_ion_box_100k()
    n = sysd["natom"]
    state = equilibrated_state(sysd)
    res = {}
    precs = {"all": ("fp64", "mixed", "fp32"), "both": ("fp64", "mixed")}.get(
        args.precision, (args.precision,))
    for prec in precs:
        sec, prof, info, drift, _rows = run_md(sysd, state, prec, args.steps, args.warmup,
                                               args.drift_steps)
        fsec, force = run_force_only(sysd, prec, 100)
        res[prec] = dict(sec=sec, prof=prof, info=info, drift=drift, fsec=fsec, force=force)
        print(f"# {prec}: {sec * 1e3:.3f} ms/MD step (force only {fsec * 1e3:.3f}), drift "
              f"{drift:.3e} eref/ns/atom, stages(ms) "
              + " ".join(f"{k}={v:.3f}" for k, v in prof.items()), file=sys.stderr, flush=True)
    head = res["mixed"] if "mixed" in res else res["fp64"]
    hp = "mixed" if "mixed" in res else "fp64"
    ncoul, nlj = pair_counts(sysd)
    flops = ncoul * (36 + 12) + nlj * 13
    t_pair = head["prof"]["pairs_timed_region"] * 1e-3
    achieved = flops / t_pair / 1e12 if t_pair > 0 else 0.0
    # the same kernel in the stage-timed run, where it runs alone (the timed
    # region runs the reciprocal chain beside it, on capped pair blocks)
    t_alone = head["prof"].get("pairs", 0.0) * 1e-3
    achieved_alone = flops / t_alone / 1e12 if t_alone > 0 else 0.0
    out = {
        "metric": "atom-steps/s (mixed SPME MD step), 100k water",
        "value": n / head["sec"],
        "unit": "atom-steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": head["sec"] * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32 (pairs, spline, FFT) + int64 fixed-point sums" if hp == "mixed" else "f64",
        "data": "synthetic",
        "config": {"workload": "100,001-atom water + Na+/Cl-, 100.1 A box, SPME 128^3 p=4, "
                               "alpha 0.32, rc 10 A, rcut_book 11 A, list every 20 steps, "
                               "NVE dt 1 fs, RATTLE",
                   "natom": n, "precision": hp},
        "ns_per_day": 86400.0 / head["sec"] * DT_FS * 1e-6,
        "force_only_ms_per_step": {p: res[p]["fsec"] * 1e3 for p in res},
        "nve_drift_eref_per_ns_per_atom": {p: res[p]["drift"] for p in res},
        "stage_ms": {p: res[p]["prof"] for p in res},
        "roofline": {"bound": "valu", "kernel": f"k_pairs ({hp})", "achieved": achieved,
                     "peak": PEAK[hp], "unit": "TFLOP/s",
                     "frac": achieved / PEAK[hp],
                     "traffic": traffic_from_profiles(f"k_pairs_{hp}"),
                     "flops_per_launch": flops, "pairs_coulomb": ncoul, "pairs_lj": nlj,
                     "avg_launch_ms": head["prof"]["pairs_timed_region"],
                     "alone": {"avg_launch_ms": t_alone * 1e3, "achieved": achieved_alone,
                               "frac": achieved_alone / PEAK[hp],
                               "run": "stage-timed run, no reciprocal chain beside it"}},
        "list": {"clusters": head["info"][0], "entries_8x4": head["info"][1],
                 "listed_pairs": head["info"][2],
                 "lane_efficiency": ncoul / (float(head["info"][3]) * head["info"][1])},
    }
    if "fp64" in res and "mixed" in res:
        f64 = res["fp64"]
        out["fp64"] = {"atom_steps_per_s": n / f64["sec"], "ms_per_step": f64["sec"] * 1e3,
                       "ns_per_day": 86400.0 / f64["sec"] * DT_FS * 1e-6,
                       "roofline_pairs": {
                           "achieved": flops / (f64["prof"]["pairs_timed_region"] * 1e-3) / 1e12,
                           "peak": PEAK["fp64"], "unit": "TFLOP/s"}}
        out["fp64"]["roofline_pairs"]["frac"] = (out["fp64"]["roofline_pairs"]["achieved"]
                                                 / PEAK["fp64"])
        out["speedup_mixed_vs_fp64"] = f64["sec"] / res["mixed"]["sec"]
    if "fp32" in res:
        f32 = res["fp32"]
        out["fp32"] = {"atom_steps_per_s": n / f32["sec"], "ms_per_step": f32["sec"] * 1e3,
                       "ns_per_day": 86400.0 / f32["sec"] * DT_FS * 1e-6}
    if args.extras:
        out["extras"] = {p: run_extras(sysd, state, p) for p in precs}
        for p, e in out["extras"].items():
            print(f"# extras {p}: MD step {e['nve_ms_per_step']:.4f} ms, with region Langevin "
                  f"{e['langevin_ms_per_step']:.4f} ms; per-atom reciprocal virial "
                  f"{e['recip_virial_atom_ms']:.3f} ms", file=sys.stderr, flush=True)
    if not args.no_heatf:
        out["heatf"] = {p: run_heatf(sysd, state, p) for p in precs if p != "fp32"}
        for p, h in out["heatf"].items():
            print(f"# heatf {p}: " + "; ".join(
                f"{k}: calforce_hf {v['calforce_hf_ms']:.2f} ms + calheatf {v['calheatf_ms']:.3f} "
                f"ms (with host arrays {v['calforce_hf_with_host_arrays_ms']:.2f} ms; fused "
                f"{v['fused_calforce_heatf_ms']:.2f} ms)"
                for k, v in h.items()), file=sys.stderr, flush=True)
    if args.slab > 0:
        sprecs = [p for p in precs if p != "fp32"]     # the slab runs FP64 and mixed
        out["slab"] = run_slab(args.slab, sprecs)
        for p in sprecs:
            e = out["slab"][p]
            print(f"# slab {p}: {out['slab']['atoms']} atoms, {e['ms_per_step']:.3f} ms/step "
                  f"({e['atom_steps_per_s']:.3e} atom-steps/s) with the chain and the heat flux "
                  f"every step", file=sys.stderr, flush=True)
    if args.interface > 0:
        iprecs = [p for p in precs if p != "fp32"]     # the interface runs FP64 and mixed
        out["interface"] = run_interface(args.interface, iprecs)
        for p in iprecs:
            e = out["interface"][p]
            print(f"# interface {p}: {out['interface']['atoms']} atoms, {e['ms_per_step']:.3f} "
                  f"ms/step ({e['atom_steps_per_s']:.3e} atom-steps/s) with both thermostats and "
                  f"300-bin heat flux every step", file=sys.stderr, flush=True)
    if reps:
        out["ensemble"] = run_ensemble(reps, args.ensemble_steps, 10, args.ensemble_hf)
        e = out["ensemble"]
        print(f"# ensemble: {e['replicas']} replicas, {e['atoms']} atoms, heat flux every "
              f"{e['heat_flux_every']} steps: sequential "
              f"{e['sequential_atom_steps_per_s']:.3e}, batched {e['batched_atom_steps_per_s']:.3e} "
              f"atom-steps/s ({e['batched_over_sequential']:.2f}x)", file=sys.stderr, flush=True)
    if not args.no_cpu:
        base, fref = cpu_leg_subprocess()
        out["cpu_baseline"] = base
        out["accuracy"] = {p: force_error(res[p]["force"], fref) for p in res}
        out["accuracy"]["against"] = f"{base['kind']} FP64 calforce on the same positions"
    print(json.dumps(out))


if __name__ == "__main__":
    main()
