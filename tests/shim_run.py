"""Drive the reference's own harness (oracle/ref_harness.F90) in a build whose
calforce, calforce_hf, mklist2a and mklist_cell are the drop-in GPU shim
(thrid_amd/fortran/calforce_gpu.F90, linked into oracle/_ref/libpgkref_gpu.so
by oracle/build_ref.sh), on the small test systems; writes the results to an
.npz.  Run as a child process of tests/test_shim.py: the harness library is a
process-wide singleton, and PGK_REF_SO must be set before pyoracle is imported.

    PGK_REF_SO=oracle/_ref/libpgkref_gpu.so PGK_PRECISION=fp64 \
        python tests/shim_run.py OUT.npz
"""
import os
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "oracle"))
sys.path.insert(0, HERE)

import numpy as np  # noqa: E402

from thrid_amd import abi  # noqa: E402
import pyoracle  # noqa: E402
from shim_cases import CASES, HF_REGIONS  # noqa: E402


def main(out):
    res = {}
    for name, (make, mode) in CASES.items():
        s = make()
        r = pyoracle.Reference(s, use_mkexcl=True)
        r.mklist(s["rcut_book"], mode)
        f = r.calforce(abi.MTS_LONG)
        for k in ("force", "ene", "virt"):
            res[f"{name}_{k}"] = f[k]
        if name in HF_REGIONS:
            z = HF_REGIONS[name](s)
            h = r.calforce_hf(dict(mode=abi.HF_VOLUME, z1=z[:-1], z2=z[1:]))
            for k in ("force", "ene", "pot_atm", "viri_atm"):
                res[f"{name}_hf_{k}"] = h[k]
        print("case", name, s["natom"], flush=True)
    np.savez(out, **res)
    print("SHIM_RUN_DONE", flush=True)


if __name__ == "__main__":
    main(sys.argv[1])
