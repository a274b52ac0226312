// k_lang.hip - the region Langevin thermostat of the reference's NVE and
// RESPA drivers on the device: registration of atoms in regions
// (langevin_regist.F90:8-364) and the two velocity half-updates
// (langevin_region.F90:8-203).
//
// Units as the MD step: positions in A, velocities in A/fs, masses in amu,
// time in fs; kT = kB T / eref * PGK_MD_CONV in amu A^2 fs^-2, so the random
// acceleration of an atom of mass m has variance 2 damp kT / (m dt) per
// component, as var_randf / m^2 with var_randf = 2 damp T m / dt in the
// reference's reduced units.
//
// Random numbers: the reference draws from SFMT (genrand_res53) in the order
// of its registration lists.  Here every draw has a fixed counter position,
// (draw * natom + atom) * 4 + k for the k-th of an atom's four uniforms in
// step `draw`, and the uniform is SplitMix64 (Steele, Lea and Flood,
// "Fast splittable pseudorandom number generators", OOPSLA 2014) evaluated
// at that position, so a thread needs no generator state and the stream
// does not depend on the order atoms are visited.  The Box-Muller transform
// is the reference's: the x and y components share sqrt(-2 ln u1) with
// cos / sin of 2 pi u2, z takes sqrt(-2 ln u3) cos(2 pi u4).
#include "pgk_device.h"

namespace {

struct LtArgs {
  int n, nreg, nunit;
  double lo[3][PGK_LT_MAXREG], hi[3][PGK_LT_MAXREG];
  double box[3];
  double vtmp[PGK_LT_MAXREG];   // 2 damp kT / dt: times m, the variance of m a
  double efac[PGK_LT_MAXREG];   // exp(-dt damp / 2)
  double h;                     // dt / 2
  uint64_t seed, ctr0;          // ctr0 = 4 natom draw
  int nstrm;                    // streaming-velocity slabs [sz0, sz1) in z
  double sz[2][PGK_LT_MAXREG], sv[3][PGK_LT_MAXREG];
  const int32_t *ustart, *uatom, *ucom;
  const double *mass;
  int32_t *reg, *sreg;
  double *frand, *det;
};

// uniform in (0, 1]: SplitMix64 output at counter position ctr
__device__ __forceinline__ double lt_uniform(uint64_t seed, uint64_t ctr) {
  uint64_t z = seed + (ctr + 1ull) * 0x9E3779B97F4A7C15ull;
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  z ^= z >> 31;
  return (double)((z >> 11) + 1ull) * 0x1.0p-53;
}

// wrapped once into [0, L) as langevin_regist.F90:96-110
__device__ __forceinline__ double lt_wrap(double g, double l) {
  if (g < 0.0) return g + l;
  if (g >= l) return g - l;
  return g;
}

__device__ __forceinline__ bool lt_in(double g, double lo, double hi, double l) {
  return (lo <= g && g < hi) || (lo <= g - l && g - l < hi);
}

// the point of registration unit u (its centre of mass or its one atom),
// each coordinate wrapped once into [0, L)
__device__ __forceinline__ void lt_unit_point(const LtArgs &a, const double *__restrict__ pos,
                                              int u, double (&g)[3]) {
#pragma clang fp contract(off)
  const int s = a.ustart[u], e = a.ustart[u + 1];
  if (a.ucom[u]) {
    double mm = 0.0, sx = 0.0, sy = 0.0, sz = 0.0;
    for (int j = s; j < e; j++) {
      const int i = a.uatom[j];
      const double m = a.mass[i];
      mm = mm + m;
      sx = sx + m * pos[3 * i];
      sy = sy + m * pos[3 * i + 1];
      sz = sz + m * pos[3 * i + 2];
    }
    g[0] = sx / mm;
    g[1] = sy / mm;
    g[2] = sz / mm;
  } else {
    const int i = a.uatom[s];
    g[0] = pos[3 * i];
    g[1] = pos[3 * i + 1];
    g[2] = pos[3 * i + 2];
  }
  for (int d = 0; d < 3; d++) g[d] = lt_wrap(g[d], a.box[d]);
}

// the streaming slab of a wrapped z (ass_strmvel.F90: z only, first match)
__device__ __forceinline__ int lt_strm_slab(const LtArgs &a, double z) {
  for (int k = 0; k < a.nstrm; k++)
    if (a.sz[0][k] <= z && z < a.sz[1][k]) return k;
  return -1;
}

// one thread per registration unit: the unit's point against the regions in
// order; the first that holds it takes every atom of the unit
// (langevin_regist.F90:112-140)
__global__ __launch_bounds__(256) void k_lt_register(LtArgs a, const double *__restrict__ pos) {
  const int u = blockIdx.x * blockDim.x + threadIdx.x;
  if (u >= a.nunit) return;
  const int s = a.ustart[u], e = a.ustart[u + 1];
  double g[3];
  lt_unit_point(a, pos, u, g);
  int r = -1;
  for (int k = 0; k < a.nreg; k++) {
    if (lt_in(g[0], a.lo[0][k], a.hi[0][k], a.box[0]) &&
        lt_in(g[1], a.lo[1][k], a.hi[1][k], a.box[1]) &&
        lt_in(g[2], a.lo[2][k], a.hi[2][k], a.box[2])) {
      r = k;
      break;
    }
  }
  const int sr = lt_strm_slab(a, g[2]);
  for (int j = s; j < e; j++) {
    a.reg[a.uatom[j]] = r;
    a.sreg[a.uatom[j]] = sr;
  }
}

// ass_strmvel (ass_strmvel.F90:8-295): every atom of a unit whose point lies
// in streaming slab k takes strmvel(k), the others zero; one thread per unit
__global__ __launch_bounds__(256) void k_strm_assign(LtArgs a, const double *__restrict__ pos,
                                                     double *__restrict__ vstrm) {
  const int u = blockIdx.x * blockDim.x + threadIdx.x;
  if (u >= a.nunit) return;
  double g[3];
  lt_unit_point(a, pos, u, g);
  const int sr = lt_strm_slab(a, g[2]);
  for (int j = a.ustart[u]; j < a.ustart[u + 1]; j++) {
    const int i = a.uatom[j];
    for (int d = 0; d < 3; d++) vstrm[3 * i + d] = sr >= 0 ? a.sv[d][sr] : 0.0;
  }
}

// One half of langevin_region.F90: FIRST draws the random accelerations of
// the registered atoms (zero for the rest) and applies the random term then
// the friction; the second half applies friction then the same random term
// to the atoms registered now.  The kinetic energy each region gains,
// sum m (v'^2 - v^2) / 2, is added to the block's own record
// det[block][region] (wave sums combined in LDS in wave order): no atomics,
// so the records, summed over blocks when read, do not depend on timing.
template <bool FIRST>
__global__ __launch_bounds__(256) void k_lt_half(LtArgs a, double *__restrict__ vel) {
#pragma clang fp contract(off)
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  const bool live = i < a.n;
  const int r = live ? a.reg[i] : -1;
  double de = 0.0;
  double fr[3] = {0.0, 0.0, 0.0};
  if (FIRST && live && r >= 0) {
    const double m = a.mass[i];
    const double var = a.vtmp[r] * m;
    const double twopi = 2.0 * 3.14159265358979323846;
    const uint64_t c = a.ctr0 + 4ull * (uint64_t)i;
    double u = lt_uniform(a.seed, c);
    fr[0] = sqrt(-2.0 * log(u) * var);
    fr[1] = sqrt(-2.0 * log(u) * var);
    u = lt_uniform(a.seed, c + 1);
    fr[0] = fr[0] * cos(twopi * u);
    fr[1] = fr[1] * sin(twopi * u);
    u = lt_uniform(a.seed, c + 2);
    fr[2] = sqrt(-2.0 * log(u) * var);
    u = lt_uniform(a.seed, c + 3);
    fr[2] = fr[2] * cos(twopi * u);
    for (int d = 0; d < 3; d++) fr[d] = fr[d] / m;
  }
  if (FIRST && live) {
    for (int d = 0; d < 3; d++) a.frand[3 * i + d] = fr[d];
  }
  // the thermostat acts on v minus the streaming velocity, which every atom
  // loses and regains (langevin_region.F90:81-82, 190)
  const int sr = live ? a.sreg[i] : -1;
  if (live && (r >= 0 || sr >= 0)) {
    double vs[3] = {0.0, 0.0, 0.0};
    if (sr >= 0)
      for (int d = 0; d < 3; d++) vs[d] = a.sv[d][sr];
    if (!FIRST && r >= 0)
      for (int d = 0; d < 3; d++) fr[d] = a.frand[3 * i + d];
    const double ef = r >= 0 ? a.efac[r] : 1.0;
    double v2o = 0.0, v2n = 0.0;
    for (int d = 0; d < 3; d++) {
      const double v0 = vel[3 * i + d] - vs[d];
      double v = v0;
      if (r >= 0) v = FIRST ? (v0 + a.h * fr[d]) * ef : v0 * ef + a.h * fr[d];
      vel[3 * i + d] = v + vs[d];
      v2o = v2o + v0 * v0;
      v2n = v2n + v * v;
    }
    if (r >= 0) de = 0.5 * a.mass[i] * (v2n - v2o) / PGK_MD_CONV;
  }
  __shared__ double sred[4][PGK_LT_MAXREG];
  for (int k = 0; k < a.nreg; k++) {
    const double t = wave_sum(r == k ? de : 0.0);
    if ((threadIdx.x & 63) == 0) sred[threadIdx.x >> 6][k] = t;
  }
  __syncthreads();
  if ((int)threadIdx.x < a.nreg) {
    const int k = threadIdx.x;
    a.det[(size_t)blockIdx.x * PGK_LT_MAXREG + k] += sred[0][k] + sred[1][k] + sred[2][k] +
                                                     sred[3][k];
  }
}

__global__ __launch_bounds__(256) void k_lt_count(int n, const int32_t *__restrict__ reg,
                                                  int32_t *__restrict__ cnt) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n && reg[i] >= 0) atomicAdd(&cnt[reg[i]], 1);
}

LtArgs lt_args(pgk_ctx *c) {
  LtArgs a{};
  a.n = c->natom;
  a.nreg = c->md_lt_nreg;
  a.nunit = c->md_lt_nunit;
  for (int d = 0; d < 3; d++) {
    a.box[d] = c->box[d];
    for (int k = 0; k < PGK_LT_MAXREG; k++) {
      a.lo[d][k] = c->md_lt_lo[d][k];
      a.hi[d][k] = c->md_lt_hi[d][k];
    }
  }
  const double dt = c->md_dt;
  for (int k = 0; k < a.nreg; k++) {
    const double kt = 1.380649e-23 / 6.9511042e-21 * c->md_lt_temp[k] * PGK_MD_CONV;
    a.vtmp[k] = 2.0 * c->md_lt_damp[k] * kt / dt;
    a.efac[k] = exp(-0.5 * dt * c->md_lt_damp[k]);
  }
  a.h = 0.5 * dt;
  a.nstrm = c->md_lt_nstrm;
  for (int k = 0; k < PGK_LT_MAXREG; k++) {
    a.sz[0][k] = c->md_lt_sz[0][k];
    a.sz[1][k] = c->md_lt_sz[1][k];
    for (int d = 0; d < 3; d++) a.sv[d][k] = c->md_lt_sv[d][k];
  }
  a.sreg = c->d_lt_sreg;
  a.seed = c->md_lt_seed;
  a.ctr0 = 4ull * (uint64_t)c->natom * c->md_lt_draw;
  a.ustart = c->d_lt_ustart;
  a.uatom = c->d_lt_uatom;
  a.ucom = c->d_lt_ucom;
  a.mass = c->d_lt_mass;
  a.reg = c->d_lt_reg;
  a.frand = c->d_lt_frand;
  a.det = c->d_lt_det;
  return a;
}

}  // namespace

int pgk_md_lt_register(pgk_ctx *c) {
  const LtArgs a = lt_args(c);
  k_lt_register<<<pgk_grid_blocks(a.nunit, 256), 256, 0, c->stream>>>(a, c->d_pos);
  HIP_CHECK(hipGetLastError());
  return PGK_OK;
}

// the atoms' streaming velocities at the current positions (no-op without
// streaming slabs)
int pgk_md_strm_assign(pgk_ctx *c) {
  if (!c->md_lt_nstrm) return PGK_OK;
  const LtArgs a = lt_args(c);
  k_strm_assign<<<pgk_grid_blocks(a.nunit, 256), 256, 0, c->stream>>>(a, c->d_pos,
                                                                     c->d_md_vstrm);
  HIP_CHECK(hipGetLastError());
  return PGK_OK;
}

// first: the start-of-step half (draws this step's accelerations and
// advances the draw counter); otherwise the end-of-step half
int pgk_md_lt_half(pgk_ctx *c, bool first) {
  const LtArgs a = lt_args(c);
  const int nb = pgk_grid_blocks(c->natom, 256);
  if (first) {
    k_lt_half<true><<<nb, 256, 0, c->stream>>>(a, c->d_md_vel);
    c->md_lt_draw++;
  } else {
    k_lt_half<false><<<nb, 256, 0, c->stream>>>(a, c->d_md_vel);
  }
  HIP_CHECK(hipGetLastError());
  return PGK_OK;
}

int pgk_md_lt_count(pgk_ctx *c) {
  HIP_CHECK(hipMemsetAsync(c->d_lt_cnt, 0, sizeof(int32_t) * PGK_LT_MAXREG, c->stream));
  k_lt_count<<<pgk_grid_blocks(c->natom, 256), 256, 0, c->stream>>>(c->natom, c->d_lt_reg,
                                                                   c->d_lt_cnt);
  HIP_CHECK(hipGetLastError());
  return PGK_OK;
}
