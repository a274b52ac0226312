"""Host side of the NVE step: constraint extraction and initial velocities."""
import numpy as np

from thrid_amd import md
from helpers import water648


def test_rigid_water_constraints():
    s = water648()
    ij, d, gs = md.rigid_constraints(s)
    assert len(gs) - 1 == 216 and len(d) == 3 * 216
    assert np.all(np.diff(gs) == 3)
    mol = s["molid"][ij - 1]
    assert np.all(mol[:, 0] == mol[:, 1])
    assert np.allclose(sorted(set(np.round(d, 6))), [1.0, 1.63299])
    # the generated geometry satisfies its own constraints
    r = np.linalg.norm(s["pos"][ij[:, 0] - 1] - s["pos"][ij[:, 1] - 1], axis=1)
    assert np.max(np.abs(r - d)) < 1e-4


def test_maxwell_boltzmann_temperature_and_momentum():
    s = water648()
    v = md.maxwell_boltzmann(s, 300.0, seed=4)
    m = s["mass"]
    assert np.allclose((m[:, None] * v).sum(0), 0.0, atol=1e-12)
    ke = 0.5 * (m[:, None] * v ** 2).sum() / md.MD_CONV
    temp = 2 * ke / (3 * len(m) * md.KB_EREF)
    assert abs(temp - 300.0) < 30.0
