"""NumPy velocity Verlet + RATTLE (Andersen 1983) driven by the oracle's forces:
the checker of the device-resident NVE step.  Same units as thrid_amd.md
(amu, A, fs, eref); constraint groups iterated until every |r|^2 is within
tol * d^2 and every |r.v_ij| within tol * d^2."""
import numpy as np

from thrid_amd import abi
from thrid_amd.md import MD_CONV


def _groups(cons_ij, group_start):
    return [np.arange(group_start[g], group_start[g + 1]) for g in range(len(group_start) - 1)]


def shake(pos, vel, old, minv, cons_ij, d2, groups, dt, tol=1e-10, maxit=500):
    for ks in groups:
        for _ in range(maxit):
            done = True
            for k in ks:
                i, j = cons_ij[k, 0] - 1, cons_ij[k, 1] - 1
                r = pos[i] - pos[j]
                ro = old[i] - old[j]
                diff = d2[k] - r @ r
                if abs(diff) <= 2 * tol * d2[k]:
                    continue
                done = False
                g = diff / (2 * (minv[i] + minv[j]) * (r @ ro))
                pos[i] += g * minv[i] * ro
                pos[j] -= g * minv[j] * ro
                vel[i] += g * minv[i] * ro / dt
                vel[j] -= g * minv[j] * ro / dt
            if done:
                break


def rattle_v(pos, vel, minv, cons_ij, d2, groups, tol=1e-10, maxit=500):
    for ks in groups:
        for _ in range(maxit):
            done = True
            for k in ks:
                i, j = cons_ij[k, 0] - 1, cons_ij[k, 1] - 1
                r = pos[i] - pos[j]
                rv = r @ (vel[i] - vel[j])
                if abs(rv) <= tol * d2[k]:
                    continue
                done = False
                kk = rv / (d2[k] * (minv[i] + minv[j]))
                vel[i] -= kk * minv[i] * r
                vel[j] += kk * minv[j] * r
            if done:
                break


def run(oracle, sysd, vel, cons_ij, cons_d, group_start, dt, nstep):
    """nstep steps with the list built once at the start (nstep < nstep_book);
    returns positions, velocities and [KE, PE, E] per step."""
    mass = np.asarray(sysd["mass"], np.float64)
    minv = 1.0 / mass
    d2 = np.asarray(cons_d) ** 2
    groups = _groups(cons_ij, group_start)
    pos = np.array(sysd["pos"], np.float64)
    vel = np.array(vel, np.float64)
    rattle_v(pos, vel, minv, cons_ij, d2, groups)
    oracle.set_positions(pos)
    oracle.mklist(sysd["rcut_book"], 0)
    r = oracle.calforce(abi.MTS_LONG)
    f = r["force"]
    rows = []

    def energies(r):
        ke = 0.5 * (mass[:, None] * vel ** 2).sum() / MD_CONV
        pe = r["ene"][abi.ENE_NAMES.index("pot_tot")]
        return [ke, pe, ke + pe]

    rows.append(energies(r))
    for _ in range(nstep):
        vel += 0.5 * dt * MD_CONV * minv[:, None] * f
        old = pos.copy()
        pos += dt * vel
        shake(pos, vel, old, minv, cons_ij, d2, groups, dt)
        oracle.set_positions(pos)
        r = oracle.calforce(abi.MTS_LONG)
        f = r["force"]
        vel += 0.5 * dt * MD_CONV * minv[:, None] * f
        rattle_v(pos, vel, minv, cons_ij, d2, groups)
        rows.append(energies(r))
    return pos, vel, np.array(rows)
