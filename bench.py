"""Benchmark: force-evaluation steps of the 100,001-atom SPME water box.

Workload (one step = one calforce call with every term on MTS_LONG and the
virial on, positions resident in HBM; the Verlet list is rebuilt every
NSTEP_BOOK steps inside the timed region): 33,333 rigid 3-site waters plus one
Na+/Cl- pair in a 100.1 A cubic box, SPME 128^3 order 4, alpha 0.32/A, 10 A
real-space and LJ cutoff, rcut_book 11 A.

Prints one JSON line.  value = atom-steps/s of the mixed form; the FP64
form's rate, the speed-up, ns/day (dt = 1 fs), the pair kernel's roofline
(HIP events on the library stream) and the reference CPU baseline are
reported beside it.
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


def run_gpu(sysd, precision, steps, warmup):
    from thrid_amd.pgk import ForceContext
    ctx = ForceContext(sysd, precision)
    ctx.mklist(sysd["rcut_book"])
    info = ctx.list_info()
    for k in range(warmup):
        if k % NSTEP_BOOK == 0:
            ctx.mklist(sysd["rcut_book"])
        e, _ = ctx.calforce_device()
    ctx.synchronize()
    ctx.profile(1)
    t0 = time.perf_counter()
    for k in range(steps):
        if k % NSTEP_BOOK == 0:
            ctx.mklist(sysd["rcut_book"])
        e, _ = ctx.calforce_device()
    ctx.synchronize()
    t1 = time.perf_counter()
    ms, calls = ctx.profile(0)
    ene = e.copy()
    ctx.close()
    stages = ["pairs", "excl", "spread", "fft", "gather", "bonded", "list", "other"]
    prof = {s: (float(ms[k] / calls[k]) if calls[k] else 0.0) for k, s in enumerate(stages)}
    return (t1 - t0) / steps, prof, ene, info


def cpu_baseline(sysd, budget_s=20.0):
    """The reference's own calforce (compiled from source, oracle/_ref) on one
    core, timed on a bounded number of calls; falls back to the C restatement."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import pyoracle
    from thrid_amd import abi
    kind = "reference" if pyoracle.reference_available() else "port"
    if kind == "reference":
        m = pyoracle.Reference(sysd)
        t0 = time.perf_counter()
        m.mklist(sysd["rcut_book"], 1)
        t_list = time.perf_counter() - t0
    else:
        m = pyoracle.Oracle(sysd)
        t0 = time.perf_counter()
        m.mklist(sysd["rcut_book"], 1)
        t_list = time.perf_counter() - t0
    ncall, t_force = 0, 0.0
    while t_force < budget_s and ncall < 50:
        t0 = time.perf_counter()
        m.calforce(abi.MTS_LONG)
        t_force += time.perf_counter() - t0
        ncall += 1
    per_step = t_force / ncall + t_list / NSTEP_BOOK
    return dict(value=sysd["natom"] / per_step, unit="atom-steps/s", cores=1, kind=kind,
                sample=f"{ncall} calforce calls + 1 list build (amortised over "
                       f"{NSTEP_BOOK} steps) of the same 100,001-atom box, serial")


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU: this build runs the force path on a single "
              "MI355X (no multi-GPU path); use --gpus 1", file=sys.stderr)
        sys.exit(2)
    from thrid_amd import systems
    sysd = systems.water_ion_box_100k()
    n = sysd["natom"]
    res = {}
    for prec in ("fp64", "mixed"):
        sec, prof, ene, info = run_gpu(sysd, prec, args.steps, args.warmup)
        res[prec] = dict(sec=sec, prof=prof, ene=ene, info=info)
        print(f"# {prec}: {sec * 1e3:.3f} ms/step, stages(ms) "
              + " ".join(f"{k}={v:.3f}" for k, v in prof.items()), file=sys.stderr, flush=True)
    ncoul, nlj = pair_counts(sysd)
    flops = ncoul * (36 + 12) + nlj * 13
    mix = res["mixed"]
    t_pair = mix["prof"]["pairs"] * 1e-3
    achieved = flops / t_pair / 1e12 if t_pair > 0 else 0.0
    out = {
        "metric": "atom-steps/s (mixed SPME force step), 100k water",
        "value": n / mix["sec"],
        "unit": "atom-steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": mix["sec"] * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32 (pairs, spline, FFT) + int64 fixed-point sums",
        "data": "synthetic",
        "config": {"workload": "100,001-atom water + Na+/Cl-, 100.1 A box, SPME 128^3 p=4, "
                               "alpha 0.32, rc 10 A, rcut_book 11 A, list every 20 steps",
                   "natom": n, "precision": "mixed"},
        "ns_per_day": 86400.0 / mix["sec"] * 1e-6,
        "fp64": {"atom_steps_per_s": n / res["fp64"]["sec"],
                 "ms_per_step": res["fp64"]["sec"] * 1e3,
                 "ns_per_day": 86400.0 / res["fp64"]["sec"] * 1e-6},
        "speedup_mixed_vs_fp64": res["fp64"]["sec"] / mix["sec"],
        "stage_ms": {p: res[p]["prof"] for p in res},
        "roofline": {"bound": "valu", "kernel": "k_pairs (mixed)", "achieved": achieved,
                     "peak": PEAK["mixed"], "unit": "TFLOP/s",
                     "frac": achieved / PEAK["mixed"], "traffic": None,
                     "flops_per_launch": flops, "pairs_coulomb": ncoul, "pairs_lj": nlj,
                     "avg_launch_ms": mix["prof"]["pairs"]},
        "list": {"clusters": res["mixed"]["info"][0], "cluster_pairs": res["mixed"]["info"][1],
                 "listed_pairs": res["mixed"]["info"][2],
                 "lane_efficiency": ncoul / (64.0 * res["mixed"]["info"][1])},
        "roofline_fp64_pairs": {"achieved": flops / (res["fp64"]["prof"]["pairs"] * 1e-3) / 1e12
                                if res["fp64"]["prof"]["pairs"] > 0 else 0.0,
                                "peak": PEAK["fp64"], "unit": "TFLOP/s"},
    }
    out["roofline_fp64_pairs"]["frac"] = out["roofline_fp64_pairs"]["achieved"] / PEAK["fp64"]
    if not args.no_cpu:
        out["cpu_baseline"] = cpu_baseline(sysd)
    print(json.dumps(out))


if __name__ == "__main__":
    main()
