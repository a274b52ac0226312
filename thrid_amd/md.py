"""Set-up helpers for the device-resident NVE step (ForceContext.md_*).

Units: masses in amu, lengths in A, time in fs, energies in eref.
"""
from __future__ import annotations

import numpy as np

# eref / (amu A^2 fs^-2) and Boltzmann's constant in eref per K
MD_CONV = 6.9511042e-21 / 1.66053907e-17
KB_EREF = 1.380649e-23 / 6.9511042e-21


def rigid_constraints(sysd):
    """Bonds with a zero force constant (the rigid water of the topology)
    become distance constraints at their reference length, grouped by
    molecule.  Returns (cons_ij 1-based (ncons, 2), cons_d, group_start)."""
    kk = np.asarray(sysd["bond_k"])
    r0 = np.asarray(sysd["bond_r0"])
    ij = np.asarray(sysd["bond_ij"]).reshape(-1, 2)
    ty = np.asarray(sysd["bond_type"]) - 1
    sel = np.abs(kk[ty]) < 1e-16
    ij, ty = ij[sel], ty[sel]
    mol = np.asarray(sysd["molid"])[ij[:, 0] - 1]
    order = np.argsort(mol, kind="stable")
    ij, ty, mol = ij[order], ty[order], mol[order]
    starts = np.flatnonzero(np.r_[True, mol[1:] != mol[:-1]])
    group_start = np.r_[starts, len(mol)].astype(np.int32)
    return ij.astype(np.int32), r0[ty].astype(np.float64), group_start


def maxwell_boltzmann(sysd, temperature, seed=1):
    """Velocities (A/fs) at the given temperature, centre-of-mass drift removed."""
    rng = np.random.default_rng(seed)
    m = np.asarray(sysd["mass"], np.float64)
    sd = np.sqrt(KB_EREF * temperature * MD_CONV / m)
    v = rng.normal(size=(len(m), 3)) * sd[:, None]
    v -= (m[:, None] * v).sum(0) / m.sum()
    return v


def drift_per_ns(samples, nsample, dt_fs, natom):
    """Least-squares slope of the total energy, in eref per ns per atom."""
    e = samples[:, 2]
    t = np.arange(len(e)) * nsample * dt_fs * 1e-6       # ns
    if len(e) < 2:
        return 0.0
    slope = np.polyfit(t, e, 1)[0]
    return float(slope / natom)


def kinetic_temperature(sysd, vel, ncons):
    """Instantaneous temperature with 3N - ncons - 3 degrees of freedom."""
    m = np.asarray(sysd["mass"], np.float64)
    ke = 0.5 * (m[:, None] * np.asarray(vel) ** 2).sum() / MD_CONV
    ndf = 3 * len(m) - ncons - 3
    return 2.0 * ke / (ndf * KB_EREF)


def equilibrate(ctx, sysd, temperature, ncycle, nstep_cycle, nstep_book, ncons):
    """Cycles of NVE steps, each followed by rescaling the velocities to the
    target temperature (a crude thermostat to relax a generated start)."""
    for _ in range(ncycle):
        ctx.md_run(nstep_cycle, nstep_book, sysd["rcut_book"], nstep_cycle)
        _p, v = ctx.md_state()
        t = kinetic_temperature(sysd, v, ncons)
        ctx.md_set_velocities(v * np.sqrt(temperature / max(t, 1e-6)))
    return ctx.md_state()


RELAX_SCHEDULE = ((0.02, 40), (0.05, 40), (0.1, 40), (0.2, 40), (0.5, 40))


def relax(ctx, sysd, cons, schedule=RELAX_SCHEDULE):
    """Quench a generated start: single MD steps from zero velocity, so each
    moves every atom by F dt^2 / 2m under the constraints (steepest descent
    with a step that grows through `schedule` = ((dt_fs, steps), ...)).
    Leaves the context set up for dt = the last schedule entry; returns the
    relaxed positions."""
    ij, d, gs = cons
    zero = np.zeros((sysd["natom"], 3))
    for dt, nstep in schedule:
        ctx.md_setup(sysd["mass"], ij, d, gs, dt)
        for _ in range(nstep):
            ctx.md_set_velocities(zero)
            ctx.md_run(1, 1, sysd["rcut_book"], 1)
    return ctx.md_state()[0]
