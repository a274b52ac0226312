"""NumPy velocity Verlet + RATTLE driven by the oracle's forces: the checker
of the device-resident NVE step.  Same units as thrid_amd.md (amu, A, fs,
eref).  The constraint stages restate the reference's: positions as rattle_c
(rattle.F90:180-226), velocities as rattlep_v (rattle_mtk.F90:455-490), each
group iterated in the given constraint order until an iteration moves nothing.
Python floats keep the reference's left-to-right, unfused evaluation."""
import numpy as np

from thrid_amd import abi
from thrid_amd.md import MD_CONV, KB_EREF


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


# ---- Nose-Hoover chain (moldyn_nhc.F90 with every term on the long step) ----

E2 = 1.0 / 6.0          # sinh series, moldyn_nhc.F90:771-774
E4 = E2 / 20.0
E6 = E4 / 42.0
E8 = E6 / 72.0


def ys_coeff(nyosh):
    """Yoshida-Suzuki weights (prepmd.F90:534-548)."""
    if nyosh == 1:
        return [1.0]
    if nyosh == 3:
        w1 = 1.0 / (2.0 - 2.0 ** (1.0 / 3.0))
        return [w1, 1.0 - 2.0 * w1, w1]
    w1 = 1.0 / (4.0 - 4.0 ** (1.0 / 3.0))
    return [w1, w1, 1.0 - 4.0 * w1, w1, w1]


class Nhc:
    """Chain state and nhcint (nhcint.F90:8-140) in this package's units:
    energies in eref, time in fs, so kT = KB_EREF T and Q in eref fs^2."""

    def __init__(self, mchain, temperature, tfreq, nfree, nyosh, nexti):
        self.m = mchain
        self.gkt = KB_EREF * temperature
        self.gnkt = self.gkt * nfree
        # prepmd.F90:515-518
        self.q = [nfree * self.gkt / tfreq / tfreq] + [self.gkt / tfreq / tfreq] * (mchain - 1)
        self.ys = ys_coeff(nyosh)
        self.next = nexti
        self.x = [0.0] * mchain
        self.v = [0.0] * mchain

    def half(self, kint, dt):
        """One nhcint call; kint = sum m v^2 in eref."""
        import math
        m, q, v, x = self.m, self.q, self.v, self.x
        wdti2 = [w * dt / self.next / 2.0 for w in self.ys]
        wdti4 = [w / 2.0 for w in wdti2]
        wdti8 = [w / 2.0 for w in wdti4]
        g = [0.0] * m
        g[0] = (kint - self.gnkt) / q[0]
        for _inc in range(self.next):
            for iys in range(len(self.ys)):
                g[m - 1] = (q[m - 2] * v[m - 2] * v[m - 2] - self.gkt) / q[m - 1]
                v[m - 1] = v[m - 1] + g[m - 1] * wdti4[iys]
                for i in range(m - 2, 0, -1):
                    g[i] = (q[i - 1] * v[i - 1] * v[i - 1] - self.gkt) / q[i]
                    e = math.exp(-1.0 * wdti8[iys] * v[i + 1])
                    v[i] = v[i] * e * e + wdti4[iys] * g[i] * e
                e = math.exp(-1.0 * wdti8[iys] * v[1])
                v[0] = v[0] * e * e + wdti4[iys] * g[0] * e
                for i in range(m):
                    x[i] = x[i] + v[i] * wdti2[iys]
                v[0] = v[0] * e * e + wdti4[iys] * g[0] * e
                for i in range(1, m - 1):
                    g[i] = (q[i - 1] * v[i - 1] * v[i - 1] - self.gkt) / q[i]
                    e = math.exp(-1.0 * wdti8[iys] * v[i + 1])
                    v[i] = v[i] * e * e + wdti4[iys] * g[i] * e
                g[m - 1] = (q[m - 2] * v[m - 2] * v[m - 2] - self.gkt) / q[m - 1]
                v[m - 1] = v[m - 1] + g[m - 1] * wdti4[iys]

    def kick_factors(self, dt):
        """expcdt2 and bbv of the thermostatted velocity update
        (moldyn_nhc.F90:1285-1300)."""
        import math
        expcdt4 = math.exp(-0.25 * dt * self.v[0])
        expcdt2 = expcdt4 * expcdt4
        argv2 = (0.25 * dt * self.v[0]) * (0.25 * dt * self.v[0])
        polyv = 1.0 + argv2 * (E2 + argv2 * (E4 + argv2 * (E6 + argv2 * E8)))
        cf = expcdt4 * polyv
        return expcdt2, cf * 0.5 * dt

    def conserved_extra(self):
        """Thermostat energy: sum Q v^2 / 2 + Nf kT x1 + kT sum x_i (i >= 2)."""
        e = sum(0.5 * q * v * v for q, v in zip(self.q, self.v))
        return e + self.gnkt * self.x[0] + self.gkt * sum(self.x[1:])


def nhc_steps(force, pos, vel, minv, dt, nstep, nhc, shake_fn=None, rattle_fn=None,
              record=None):
    """nstep NVT steps in the order of moldyn_nhc.F90 (nhcint, thermostatted
    long-force velocity update, drift, rattle_c, forces, velocity update,
    rattlep_v, nhcint) with one short step per long step.  force(pos) returns
    (F in eref/A, PE in eref); shake_fn(pos, vel, old) and rattle_fn(pos, vel)
    apply the constraints in place; record(pos, vel, pe) is called at step 0
    and after every step.  Returns the final (pos, vel)."""
    mass = 1.0 / minv
    f, pe = force(pos)
    if record:
        record(pos, vel, pe)
    for _ in range(nstep):
        nhc.half((mass[:, None] * vel * vel).sum() / MD_CONV, dt)
        s2, bbv = nhc.kick_factors(dt)
        acc = f * MD_CONV * minv[:, None]
        vel = vel * s2 + acc * bbv
        old = pos.copy()
        pos = pos + dt * vel
        if shake_fn:
            shake_fn(pos, vel, old)
        f, pe = force(pos)
        acc = f * MD_CONV * minv[:, None]
        vel = vel * s2 + acc * bbv
        if rattle_fn:
            rattle_fn(pos, vel)
        nhc.half((mass[:, None] * vel * vel).sum() / MD_CONV, dt)
        if record:
            record(pos, vel, pe)
    return pos, vel


def run_nhc(oracle, sysd, vel, cons_ij, cons_d, group_start, dt, nstep, nhc, record=None):
    """nhc_steps on the oracle's forces with the rigid-water constraints
    (the list built once at the start, nstep < nstep_book)."""
    mass = np.asarray(sysd["mass"], np.float64)
    minv = 1.0 / mass
    d2 = np.asarray(cons_d) ** 2
    groups = _groups(cons_ij, group_start)
    pos = np.array(sysd["pos"], np.float64)
    vel = np.array(vel, np.float64)
    rattle_v(pos, vel, minv, cons_ij, d2, groups)
    oracle.set_positions(pos)
    oracle.mklist(sysd["rcut_book"], 0)

    def force(p):
        oracle.set_positions(p)
        r = oracle.calforce(abi.MTS_LONG)
        return r["force"], r["ene"][abi.ENE_NAMES.index("pot_tot")]

    return nhc_steps(force, pos, vel, minv, dt, nstep, nhc,
                     lambda p, v, o: shake(p, v, o, minv, cons_ij, d2, groups, dt),
                     lambda p, v: rattle_v(p, v, minv, cons_ij, d2, groups), record)


def run_respa(oracle, sysd, vel, cons_ij, cons_d, group_start, dt_long, nshort, nstep, mts):
    """RESPA steps as moldyn.F90:1465-1771: nshort inner steps of
    dt_long / nshort; the long half kick before the first inner kick and
    after the last; long forces after the last inner drift; rattle_c at the
    inner step, rattlep_v after every inner step.  Returns positions,
    velocities and [KE, PE, E] per outer step (PE: every term once)."""
    mass = np.asarray(sysd["mass"], np.float64)
    minv = 1.0 / mass
    d2 = np.asarray(cons_d) ** 2
    groups = _groups(cons_ij, group_start)
    pos = np.array(sysd["pos"], np.float64)
    vel = np.array(vel, np.float64)
    rattle_v(pos, vel, minv, cons_ij, d2, groups)
    oracle.set_positions(pos)
    oracle.mklist(sysd["rcut_book"], 0)
    itot = abi.ENE_NAMES.index("pot_tot")
    rs = oracle.calforce(abi.MTS_SHORT, mts)
    rl = oracle.calforce(abi.MTS_LONG, mts)
    fs, fl = rs["force"], rl["force"]
    dts = dt_long / nshort

    iewc = abi.ENE_NAMES.index("pot_ewc")

    def row():
        # each call's pot_tot holds the terms it evaluated plus the Ewald self
        # energy (a constant), which the step counts once
        ke = 0.5 * (mass[:, None] * vel ** 2).sum() / MD_CONV
        pe = rl["ene"][itot] + (rs["ene"][itot] - rs["ene"][iewc])
        return [ke, pe, ke + pe]

    rows = [row()]
    for _ in range(nstep):
        for s in range(1, nshort + 1):
            if s == 1:
                vel += 0.5 * dt_long * MD_CONV * minv[:, None] * fl
            vel += 0.5 * dts * MD_CONV * minv[:, None] * fs
            old = pos.copy()
            pos += dts * vel
            shake(pos, vel, old, minv, cons_ij, d2, groups, dts)
            oracle.set_positions(pos)
            rs = oracle.calforce(abi.MTS_SHORT, mts)
            fs = rs["force"]
            if s == nshort:
                rl = oracle.calforce(abi.MTS_LONG, mts)
                fl = rl["force"]
            vel += 0.5 * dts * MD_CONV * minv[:, None] * fs
            if s == nshort:
                vel += 0.5 * dt_long * MD_CONV * minv[:, None] * fl
            rattle_v(pos, vel, minv, cons_ij, d2, groups)
        rows.append(row())
    return pos, vel, np.array(rows)
