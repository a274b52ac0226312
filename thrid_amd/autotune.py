"""Autotuner of the pair kernel's launch shape.

The pair kernel has two forms (pgk_tune, PGK_TUNE_PAIR_CHUNK_BELOW): one wave
per 8-atom i-cluster walking the cluster's whole list, or the chunked form,
one wave per 64 list entries of a cluster.  The first wastes nothing on large
systems; the second fills the device when a system has too few clusters for
one wave each (the 1.5k-12k-atom replicas of a heat-flux ensemble, the
18k-atom slab).  This module times
both forms on a ladder of water boxes, takes the crossover in cluster count,
and writes it with the measurements to tuning.json beside this file, which
ForceContext applies at creation (the library's compiled default,
PGK_TUNE_DEFAULT_CHUNK_BELOW, is the committed table's value).

Run on the GPU:  python -m thrid_amd.autotune [--write]
"""
import json
import math
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
TABLE = os.path.join(HERE, "tuning.json")
LADDER = (1500, 3000, 6000, 12000, 24000, 48000, 100000)


def load_table():
    """The committed tuning table, or None."""
    try:
        with open(TABLE) as f:
            return json.load(f)
    except (OSError, ValueError):
        return None


def _pair_us(ctx, reps=20):
    """Mean device time of the pair kernel over reps force calls."""
    for _ in range(3):
        ctx.calforce_async()
    ctx.synchronize()
    ctx.profile_stages(1)
    ctx.profile(1)
    for _ in range(reps):
        ctx.calforce_async()
    ctx.synchronize()
    ms, calls = ctx.profile(0)
    ctx.profile_stages(0xFFFFFFFF)
    return float(ms[0] / calls[0] * 1e3)


def measure(precisions=("mixed", "fp64"), ladder=LADDER):
    """Both forms of the pair kernel on water boxes of the ladder's sizes."""
    from thrid_amd import abi, systems
    from thrid_amd.pgk import ForceContext
    rows = []
    for n in ladder:
        s = systems.mixture_box(n, "water", seed=11)
        for prec in precisions:
            c = ForceContext(s, prec, tuned=False)
            c.mklist()
            ncl = c.list_info()[0]
            c.tune(abi.TUNE_PAIR_CHUNK_BELOW, 0)
            whole = _pair_us(c)
            c.tune(abi.TUNE_PAIR_CHUNK_BELOW, 1 << 40)
            chunked = _pair_us(c)
            c.close()
            rows.append({"natom": int(s["natom"]), "precision": prec, "clusters": int(ncl),
                         "per_cluster_us": whole, "chunked_us": chunked})
            print(f"{prec} {s['natom']} atoms, {ncl} clusters: per cluster {whole:.1f} us, "
                  f"chunked {chunked:.1f} us", flush=True)
    return rows


def crossover(rows):
    """The cluster count below which the chunked form wins: between the
    largest system where it wins in every precision and the next one (their
    geometric mean, rounded)."""
    by_ncl = {}
    for r in rows:
        by_ncl.setdefault(r["clusters"], []).append(r["chunked_us"] < r["per_cluster_us"])
    ncls = sorted(by_ncl)
    wins = [all(by_ncl[k]) for k in ncls]
    last = -1
    for i, w in enumerate(wins):
        if not w:
            break
        last = i
    if last < 0:
        return 0
    if last == len(ncls) - 1:
        return int(2 * ncls[-1])
    return int(round(math.sqrt(ncls[last] * ncls[last + 1])))


def main(argv):
    import torch
    rows = measure()
    below = crossover(rows)
    props = torch.cuda.get_device_properties(0)
    table = {"pair_chunk_below": below,
             "device": f"{props.name} ({props.gcnArchName}, {props.multi_processor_count} CUs)",
             "ladder": "water boxes (systems.mixture_box(n, 'water')), rc 10 A, rcut_book 11 A",
             "measurements": rows}
    print(json.dumps({"pair_chunk_below": below}))
    if "--write" in argv:
        out = os.environ.get("PGK_TUNING_OUT", TABLE)
        with open(out, "w") as f:
            json.dump(table, f, indent=1)
    return 0


if __name__ == "__main__":
    sys.exit(main(sys.argv[1:]))
