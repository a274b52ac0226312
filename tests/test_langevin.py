"""The NumPy restatement of the region Langevin thermostat
(langevin_regist.F90:8-364, langevin_region.F90:8-203), pinned by what fixes
it independently: SplitMix64's published outputs, the moments of the
Box-Muller accelerations, hand-worked registrations (wrapping, the one-period
reach of a region, first match, centre of mass against atom by atom), the
closed-form stationary velocity variance of free particles and the energy
record against the kinetic energy change."""
import math

import numpy as np

import md_reference as mr
from thrid_amd.md import MD_CONV, KB_EREF


def test_splitmix64_known_outputs():
    # the first outputs of SplitMix64 seeded with 0 (the generator's
    # reference sequence, e.g. the seeding vectors of the xoshiro family)
    z = mr.splitmix64(0, np.arange(3, dtype=np.uint64))
    assert [int(v) for v in z] == [0xE220A8397B1DCDAF, 0x6E789E6AA1B965F4, 0x06C45D188009454F]
    u = mr.uniform(0, np.arange(3, dtype=np.uint64))
    assert np.all(u == ((z >> np.uint64(11)) + np.uint64(1)).astype(np.float64) * 2.0 ** -53)


def test_uniforms_in_half_open_interval():
    u = mr.uniform(987654321, np.arange(2_000_000, dtype=np.uint64))
    assert u.min() > 0.0 and u.max() <= 1.0
    assert abs(u.mean() - 0.5) < 1e-3
    assert abs(u.var() * 12.0 - 1.0) < 3e-3


def _free(n, mass=16.0, box=(30.0, 30.0, 30.0)):
    molid = np.arange(1, n + 1)
    return np.full(n, mass), molid, box


def test_random_accelerations_are_standard_normal():
    """From rest, the first half leaves v = dt/2 a exp(-dt damp/2): a scaled
    back by its standard deviation has zero mean, unit variance, x and y
    uncorrelated (cos / sin of the same angle) and normal kurtosis."""
    n = 200_000
    mass, molid, box = _free(n)
    dt, damp, temp = 2.0, 0.1, 300.0
    lt = mr.Langevin([[-1e9, 1e9, -1e9, 1e9, -1e9, 1e9]], temp, damp, molid, mass, box, dt,
                     seed=5)
    pos = np.random.default_rng(1).uniform(0, 30, size=(n, 3))
    lt.register(pos)
    assert np.all(lt.reg == 0)
    vel = lt.half(np.zeros((n, 3)), True)
    sd = math.sqrt(2.0 * damp * KB_EREF * temp * MD_CONV / (mass[0] * dt))
    z = vel / (0.5 * dt * math.exp(-0.5 * dt * damp) * sd)
    assert np.all(np.abs(z.mean(axis=0)) < 0.01)
    assert np.all(np.abs(z.var(axis=0) - 1.0) < 0.01)
    assert abs(np.corrcoef(z[:, 0], z[:, 1])[0, 1]) < 0.01
    assert np.all(np.abs((z ** 4).mean(axis=0) - 3.0) < 0.05)


def test_registration_hand_cases():
    box = (10.0, 10.0, 10.0)
    # region 0: x in [-2, 2) (reaches x >= 8 through the period), every y, z;
    # region 1: z in [5, 10), every x, y.  Atoms are their own molecules.
    regions = [[-2.0, 2.0, 0.0, 10.0, 0.0, 10.0], [0.0, 10.0, 0.0, 10.0, 5.0, 10.0]]
    pos = np.array([
        [9.0, 1.0, 1.0],     # 9 - 10 = -1 in [-2, 2)            -> 0
        [-1.0, 1.0, 1.0],    # wrapped to 9, then as above       -> 0
        [3.0, 1.0, 7.0],     # only region 1                     -> 1
        [1.0, 1.0, 7.0],     # both: the first wins              -> 0
        [3.0, 1.0, 2.0],     # neither                           -> -1
        [10.5, 1.0, 1.0],    # wrapped to 0.5                    -> 0
        [5.0, 1.0, 10.0],    # z = 10 wraps to 0: neither        -> -1
        [5.0, 1.0, -0.5],    # z wraps to 9.5                    -> 1
        [2.0, 1.0, 4.999],   # x = 2 is outside [-2, 2)          -> -1
    ])
    n = len(pos)
    lt = mr.Langevin(regions, 300.0, 0.01, np.arange(1, n + 1), np.ones(n), box, 1.0)
    assert list(lt.register(pos)) == [0, 0, 1, 0, -1, 0, -1, 1, -1]


def test_registration_by_centre_of_mass():
    """A two-atom molecule straddling x = 2: by its centre of mass (x = 2.0,
    outside [-2, 2)) neither atom is registered; atom by atom the first is."""
    box = (10.0, 10.0, 10.0)
    regions = [[-2.0, 2.0, 0.0, 10.0, 0.0, 10.0]]
    pos = np.array([[1.5, 1.0, 1.0], [2.5, 1.0, 1.0], [0.5, 1.0, 1.0]])
    molid = np.array([1, 1, 2])
    mass = np.array([1.0, 1.0, 3.0])
    com = mr.Langevin(regions, 300.0, 0.01, molid, mass, box, 1.0)
    assert list(com.register(pos)) == [-1, -1, 0]
    atom = mr.Langevin(regions, 300.0, 0.01, molid, mass, box, 1.0, mol_com=[0, 0])
    assert list(atom.register(pos)) == [0, -1, 0]
    # heavier first atom pulls the centre inside
    heavy = mr.Langevin(regions, 300.0, 0.01, molid, np.array([3.0, 1.0, 3.0]), box, 1.0)
    assert list(heavy.register(pos)) == [0, 0, 0]


def test_free_particle_stationary_variance_and_energy_record():
    """Free particles (no force between the halves) relax to the closed-form
    variance (kT/m)(x coth x), x = damp dt / 2 -- 8 % above kT/m here, so
    a half applied in the wrong order or with the wrong factor fails -- and
    the energy record equals the kinetic energy gained."""
    n = 40_000
    mass, molid, box = _free(n, mass=12.0)
    dt, damp, temp = 1.0, 1.0, 250.0
    lt = mr.Langevin([[-1e9, 1e9, -1e9, 1e9, -1e9, 1e9]], temp, damp, molid, mass, box, dt,
                     seed=11)
    lt.register(np.zeros((n, 3)))
    vel = np.zeros((n, 3))
    var = []
    for s in range(40):
        lt.half(vel, True)
        lt.half(vel, False)
        if s >= 20:
            var.append(vel.var())
    kt_m = KB_EREF * temp * MD_CONV / 12.0
    want = mr.langevin_variance(kt_m, damp, dt)
    assert abs(want / kt_m - 1.0820) < 1e-3
    assert abs(np.mean(var) / want - 1.0) < 0.01
    ke = 0.5 * (mass[:, None] * vel * vel).sum() / MD_CONV
    assert abs(lt.det.sum() - ke) < 1e-9 * ke


def test_outside_atoms_untouched_and_joiners_get_friction_only():
    """Atoms outside every region keep their velocity; an atom that joins a
    region between the halves gets the friction factor and no random term
    (its acceleration was zeroed at the first half)."""
    box = (10.0, 10.0, 10.0)
    regions = [[0.0, 5.0, 0.0, 10.0, 0.0, 10.0]]
    pos = np.array([[1.0, 1.0, 1.0], [7.0, 1.0, 1.0]])
    lt = mr.Langevin(regions, 300.0, 0.1, np.array([1, 2]), np.array([16.0, 16.0]), box, 1.0,
                     seed=3)
    lt.register(pos)
    v = np.array([[0.01, 0.0, 0.0], [0.0, 0.02, 0.0]])
    lt.half(v, True)
    assert np.all(v[1] == [0.0, 0.02, 0.0])
    pos[1, 0] = 4.0                        # atom 2 moved into the region
    lt.register(pos)
    lt.half(v, False)
    assert np.all(v[1] == np.array([0.0, 0.02, 0.0]) * math.exp(-0.05))


def test_streaming_velocity():
    """With streaming slabs (ass_strmvel.F90: wrapped z of the unit, first
    slab, no reach through the period) the thermostat acts on v - v_s: at
    T = 0 the peculiar velocity decays by exp(-dt damp / 2) per half while
    v_s is kept; atoms in no thermostat region keep their velocity."""
    box = (10.0, 10.0, 10.0)
    regions = [[0.0, 10.0, 0.0, 10.0, 0.0, 5.0]]
    stream = ([0.0, 2.0], [2.0, 8.0], [[0.01, 0.0, 0.0], [0.0, -0.02, 0.0]])
    pos = np.array([[1.0, 1.0, 1.0],    # region 0, slab 0
                    [1.0, 1.0, 3.0],    # region 0, slab 1
                    [1.0, 1.0, 7.0],    # no region, slab 1
                    [1.0, 1.0, -1.5],   # z wraps to 8.5: no region, no slab
                    [1.0, 1.0, 4.0]])   # region 0, slab 1
    n = len(pos)
    lt = mr.Langevin(regions, 0.0, 0.2, np.arange(1, n + 1), np.full(n, 16.0), box, 1.0,
                     stream=stream)
    lt.register(pos)
    assert list(lt.reg) == [0, 0, -1, -1, 0]
    np.testing.assert_array_equal(lt.vstrm[:, 0], [0.01, 0.0, 0.0, 0.0, 0.0])
    np.testing.assert_array_equal(lt.vstrm[:, 1], [0.0, -0.02, -0.02, 0.0, -0.02])
    v0 = np.array([[0.03, 0.0, 0.0], [0.0, 0.01, 0.0], [0.0, 0.05, 0.0], [0.0, 0.0, 0.04],
                   [0.02, -0.02, 0.0]])
    v = v0.copy()
    lt.half(v, True)
    f = math.exp(-0.1)
    want = lt.vstrm + (v0 - lt.vstrm) * f
    want[2:4] = v0[2:4]
    np.testing.assert_allclose(v, want, rtol=0, atol=1e-17)
    np.testing.assert_array_equal(v[4], [0.02 * f, -0.02, 0.0])   # at its stream in y
