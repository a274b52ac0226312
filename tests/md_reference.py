"""NumPy velocity Verlet + RATTLE driven by the oracle's forces: the checker
of the device-resident NVE step.  Same units as thrid_amd.md (amu, A, fs,
eref).  The constraint stages restate the reference's: positions as rattle_c
(rattle.F90:180-226), velocities as rattlep_v (rattle_mtk.F90:455-490), each
group iterated in the given constraint order until an iteration moves nothing.
Python floats keep the reference's left-to-right, unfused evaluation."""
import numpy as np

from thrid_amd import abi
from thrid_amd.md import MD_CONV


def _groups(cons_ij, group_start):
    return [np.arange(group_start[g], group_start[g + 1]) for g in range(len(group_start) - 1)]


def _dot(a, b):
    return a[0] * b[0] + a[1] * b[1] + a[2] * b[2]


def shake(pos, vel, old, minv, cons_ij, d2, groups, dt, eps=1e-10, maxit=500):
    for ks in groups:
        for _ in range(maxit):
            done = True
            for k in ks:
                i, j = cons_ij[k, 0] - 1, cons_ij[k, 1] - 1
                r = [pos[i, c] - pos[j, c] for c in range(3)]
                rr2 = _dot(r, r)
                if abs(rr2 - d2[k]) / d2[k] < eps:
                    continue
                done = False
                ro = [old[i, c] - old[j, c] for c in range(3)]
                g = (rr2 - d2[k]) / (2.0 * _dot(r, ro) * (minv[i] + minv[j]) * dt)
                for c in range(3):
                    gr = g * ro[c]
                    pos[i, c] = pos[i, c] - minv[i] * gr * dt
                    pos[j, c] = pos[j, c] + minv[j] * gr * dt
                    vel[i, c] = vel[i, c] - minv[i] * gr
                    vel[j, c] = vel[j, c] + minv[j] * gr
            if done:
                break
        else:
            raise RuntimeError("rattle_c did not converge")


def rattle_v(pos, vel, minv, cons_ij, d2, groups, eps=1e-10, maxit=500):
    eps2 = eps * eps
    for ks in groups:
        for _ in range(maxit):
            done = True
            for k in ks:
                i, j = cons_ij[k, 0] - 1, cons_ij[k, 1] - 1
                r = [pos[i, c] - pos[j, c] for c in range(3)]
                v = [vel[i, c] - vel[j, c] for c in range(3)]
                rv = _dot(r, v)
                if rv * rv / _dot(r, r) * _dot(v, v) < eps2:
                    continue
                done = False
                kk = rv / (d2[k] * (minv[i] + minv[j]))
                for c in range(3):
                    kr = kk * r[c]
                    vel[i, c] = vel[i, c] - minv[i] * kr
                    vel[j, c] = vel[j, c] + minv[j] * kr
            if done:
                break
        else:
            raise RuntimeError("rattlep_v did not converge")


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
