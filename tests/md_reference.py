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


def run(oracle, sysd, vel, cons_ij, cons_d, group_start, dt, nstep, lang=None):
    """nstep steps with the list built once at the start (nstep < nstep_book);
    returns positions, velocities and [KE, PE, E] per step.  lang (a
    Langevin) adds the region thermostat's halves around each step."""
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
        # kinetic energy relative to the streaming velocities (calkin.F90:98)
        u = vel - lang.vstrm if lang else vel
        ke = 0.5 * (mass[:, None] * u ** 2).sum() / MD_CONV
        pe = r["ene"][abi.ENE_NAMES.index("pot_tot")]
        return [ke, pe, ke + pe]

    if lang:
        lang.register(pos)
    rows.append(energies(r))
    for _ in range(nstep):
        if lang:
            lang.half(vel, True)
        vel += 0.5 * dt * MD_CONV * minv[:, None] * f
        old = pos.copy()
        pos += dt * vel
        shake(pos, vel, old, minv, cons_ij, d2, groups, dt)
        oracle.set_positions(pos)
        r = oracle.calforce(abi.MTS_LONG)
        f = r["force"]
        vel += 0.5 * dt * MD_CONV * minv[:, None] * f
        rattle_v(pos, vel, minv, cons_ij, d2, groups)
        if lang:
            lang.register(pos)
            lang.half(vel, False)
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
              record=None, lang=None, strm=None):
    """nstep NVT steps in the order of moldyn_nhc.F90 (nhcint, thermostatted
    long-force velocity update, drift, rattle_c, streaming velocities,
    forces, velocity update, rattlep_v, nhcint) with one short step per long
    step.  force(pos) returns (F in eref/A, PE in eref); shake_fn(pos, vel,
    old) and rattle_fn(pos, vel) apply the constraints in place; record(pos,
    vel, pe) is called at step 0 and after every step.  strm: an object whose
    register(pos) assigns its per-atom streaming velocities .vstrm
    (ass_strmvel.F90, e.g. a Langevin with streaming slabs): the chain's
    kinetic energy and half kicks then act on v - vstrm
    (moldyn_nhc.F90:1280-1310, 1412, nhcint.F90:53-54).  Returns the final
    (pos, vel)."""
    mass = 1.0 / minv
    f, pe = force(pos)
    if strm:
        strm.register(pos)
    if record:
        record(pos, vel, pe)
    # with a region Langevin thermostat (lang) the chain acts on the atoms
    # outside every region only; the others take plain half kicks
    chain = np.ones(len(mass), bool)
    if lang:
        lang.register(pos)

    def vs():
        return strm.vstrm if strm else 0.0

    def kint():
        u = (vel - vs())[chain]
        return (mass[chain, None] * u * u).sum() / MD_CONV

    def kick(vel, f, s2, bbv):
        acc = f * MD_CONV * minv[:, None]
        out = ((vel - vs()) * s2 + acc * bbv) + vs() if strm else vel * s2 + acc * bbv
        if lang:
            plain = vel + 0.5 * dt * MD_CONV * minv[:, None] * f
            out[~chain] = plain[~chain]
        return out

    for _ in range(nstep):
        if lang:
            lang.half(vel, True)
            chain = lang.reg < 0
        nhc.half(kint(), dt)
        s2, bbv = nhc.kick_factors(dt)
        vel = kick(vel, f, s2, bbv)
        old = pos.copy()
        pos = pos + dt * vel
        if shake_fn:
            shake_fn(pos, vel, old)
        if strm:
            strm.register(pos)
        f, pe = force(pos)
        vel = kick(vel, f, s2, bbv)
        if rattle_fn:
            rattle_fn(pos, vel)
        nhc.half(kint(), dt)
        if lang:
            lang.register(pos)
            lang.half(vel, False)
        if record:
            record(pos, vel, pe)
    return pos, vel


def run_nhc(oracle, sysd, vel, cons_ij, cons_d, group_start, dt, nstep, nhc, record=None,
            lang=None, strm=None):
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
                     lambda p, v: rattle_v(p, v, minv, cons_ij, d2, groups), record, lang,
                     strm)


def run_respa(oracle, sysd, vel, cons_ij, cons_d, group_start, dt_long, nshort, nstep, mts,
              lang=None):
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
        u = vel - lang.vstrm if lang else vel
        ke = 0.5 * (mass[:, None] * u ** 2).sum() / MD_CONV
        pe = rl["ene"][itot] + (rs["ene"][itot] - rs["ene"][iewc])
        return [ke, pe, ke + pe]

    if lang:
        lang.register(pos)
    rows = [row()]
    for _ in range(nstep):
        if lang:
            lang.half(vel, True)
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
        if lang:
            lang.register(pos)
            lang.half(vel, False)
        rows.append(row())
    return pos, vel, np.array(rows)


# ---- region Langevin thermostat (langevin_regist.F90, langevin_region.F90) ----

_SM_GAMMA = np.uint64(0x9E3779B97F4A7C15)


def splitmix64(seed, ctr):
    """SplitMix64 (Steele, Lea and Flood, OOPSLA 2014) output at counter
    positions ctr: the (ctr+1)-th output of the generator seeded with seed."""
    z = np.uint64(seed) + (np.asarray(ctr, np.uint64) + np.uint64(1)) * _SM_GAMMA
    z = (z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
    z = (z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
    return z ^ (z >> np.uint64(31))


def uniform(seed, ctr):
    """The top 53 bits of splitmix64 mapped to (0, 1]."""
    return ((splitmix64(seed, ctr) >> np.uint64(11)) + np.uint64(1)).astype(np.float64) \
        * 2.0 ** -53


class Langevin:
    """The region Langevin thermostat in this package's units (A, fs, amu):
    registration as langevin_regist.F90:8-364 (first region holding the
    wrapped point; a molecule flagged in mol_com by its centre of mass, the
    others atom by atom) and the two halves of langevin_region.F90:8-203
    (streaming velocity zero), with the counter-based uniforms of
    splitmix64 in place of SFMT: atom i's k-th uniform in draw d sits at
    counter (d natom + i) 4 + k."""

    def __init__(self, regions, temperature, damp, molid, mass, box, dt, mol_com=None, seed=0,
                 stream=None):
        import math
        rb = np.asarray(regions, np.float64)
        self.reg_box = rb.reshape(0, 6) if rb.size == 0 else np.atleast_2d(rb)
        nr = self.reg_box.shape[0]
        self.temp = np.broadcast_to(np.asarray(temperature, np.float64), (nr,))
        self.damp = np.broadcast_to(np.asarray(damp, np.float64), (nr,))
        self.mass = np.asarray(mass, np.float64)
        self.n = len(self.mass)
        self.box = [float(b) for b in box]
        self.dt = dt
        self.seed = seed
        self.draw = 0
        molid = np.asarray(molid)
        nmol = int(molid.max())
        self.units = []          # (com?, [atoms])
        for m in range(nmol):
            atoms = [int(i) for i in np.nonzero(molid == m + 1)[0]]
            if not atoms:
                continue
            if mol_com is None or mol_com[m]:
                self.units.append((True, atoms))
            else:
                self.units += [(False, [i]) for i in atoms]
        kt = [1.380649e-23 / 6.9511042e-21 * float(t) * MD_CONV for t in self.temp]
        self.vtmp = [2.0 * float(g) * k / dt for g, k in zip(self.damp, kt)]
        self.efac = [math.exp(-0.5 * dt * float(g)) for g in self.damp]
        self.reg = np.full(self.n, -1, np.int64)
        # streaming-velocity slabs (ass_strmvel.F90): z ranges and velocities
        self.stream = None if stream is None else (
            [float(z) for z in stream[0]], [float(z) for z in stream[1]],
            np.asarray(stream[2], np.float64).reshape(-1, 3))
        self.vstrm = np.zeros((self.n, 3))
        self.frand = np.zeros((self.n, 3))
        self.det = np.zeros(nr)

    def region_of(self, g):
        """First region holding point g (already wrapped), or -1
        (langevin_regist.F90:112-140: each range may reach one period past
        the cell)."""
        for k, (x1, x2, y1, y2, z1, z2) in enumerate(self.reg_box):
            ok = True
            for d, (lo, hi) in enumerate(((x1, x2), (y1, y2), (z1, z2))):
                l = self.box[d]
                if not ((lo <= g[d] < hi) or (lo <= g[d] - l < hi)):
                    ok = False
                    break
            if ok:
                return k
        return -1

    def register(self, pos):
        for com, atoms in self.units:
            if com:
                mm, s = 0.0, [0.0, 0.0, 0.0]
                for i in atoms:
                    m = float(self.mass[i])
                    mm = mm + m
                    for d in range(3):
                        s[d] = s[d] + m * float(pos[i, d])
                g = [s[d] / mm for d in range(3)]
            else:
                g = [float(pos[atoms[0], d]) for d in range(3)]
            for d in range(3):
                if g[d] < 0.0:
                    g[d] = g[d] + self.box[d]
                elif g[d] >= self.box[d]:
                    g[d] = g[d] - self.box[d]
            r = self.region_of(g)
            vs = np.zeros(3)
            if self.stream is not None:
                for k, (z1, z2) in enumerate(zip(self.stream[0], self.stream[1])):
                    if z1 <= g[2] < z2:
                        vs = self.stream[2][k]
                        break
            for i in atoms:
                self.reg[i] = r
                self.vstrm[i] = vs
        return self.reg

    def half(self, vel, first):
        """One half of langevin_region.F90 on vel (in place): every atom
        loses its streaming velocity first and regains it last."""
        vel -= self.vstrm
        on = self.reg >= 0
        idx = np.nonzero(on)[0]
        r = self.reg[idx]
        h = 0.5 * self.dt
        if first:
            c = (np.uint64(self.draw) * np.uint64(self.n) + idx.astype(np.uint64)) * np.uint64(4)
            m = self.mass[idx]
            var = np.array([self.vtmp[k] for k in r]) * m
            twopi = 2.0 * np.pi
            u = uniform(self.seed, c)
            fx = np.sqrt(-2.0 * np.log(u) * var)
            fy = np.sqrt(-2.0 * np.log(u) * var)
            u = uniform(self.seed, c + np.uint64(1))
            fx = fx * np.cos(twopi * u)
            fy = fy * np.sin(twopi * u)
            u = uniform(self.seed, c + np.uint64(2))
            fz = np.sqrt(-2.0 * np.log(u) * var)
            u = uniform(self.seed, c + np.uint64(3))
            fz = fz * np.cos(twopi * u)
            self.frand[:] = 0.0
            self.frand[idx] = np.stack([fx / m, fy / m, fz / m], axis=1)
            self.draw += 1
        fr = self.frand[idx]
        ef = np.array([self.efac[k] for k in r])[:, None]
        v0 = vel[idx]
        v = (v0 + h * fr) * ef if first else v0 * ef + h * fr
        vel[idx] = v
        de = 0.5 * self.mass[idx] * ((v * v).sum(axis=1) - (v0 * v0).sum(axis=1)) / MD_CONV
        np.add.at(self.det, r, de)
        vel += self.vstrm
        return vel


def langevin_variance(kt_over_m, damp, dt):
    """Stationary per-component velocity variance of a free particle under
    the two halves of langevin_region.F90 with no force in between:
    v' = v c + (dt/2) a (1 + c), c = exp(-damp dt), a ~ N(0, 2 damp kT /
    (m dt)), so var = (kT/m) (damp dt / 2) coth(damp dt / 2)."""
    x = 0.5 * damp * dt
    return kt_over_m * x / np.tanh(x)
