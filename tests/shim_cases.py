"""Systems the drop-in shim is checked on (tests/test_shim.py, tests/shim_run.py):
name -> (builder, reference list mode), and the heat-flux slab edges of the
cases that also run calforce_hf."""
import numpy as np

from helpers import water648, small_slab

CASES = {
    "water648": (water648, 0),            # mklist2a (fewer than 3 cells per axis)
    "slab_rb": (lambda: small_slab("rb"), 1),   # mklist_cell, bonded + RB torsions + 1-4
}
HF_REGIONS = {
    "slab_rb": lambda s: np.linspace(0.0, float(s["box"][2]), 7),
}
