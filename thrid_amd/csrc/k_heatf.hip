// k_heatf.hip - calheatf (calheatf.F90:8-1470) on the device: the heat flux
// of each z region from the per-atom energies and virial arrays that the
// heat-flux pass of calforce_hf leaves in HBM (k_hf.hip, pgk_hf.h layout).
//
// Per region r, with the centre of atom i (the atom, or its molecule's
// centre of mass for hftyp MOLE, z wrapped once into [0, Lz)) and the
// velocity of that centre (the atom's, or the molecule's mass-weighted mean):
//   transport   sum_{z1 <= z_i <= z2} (m_i u_i^2 / 2) vc_i   and   e_c(i) vc_i
//   collision   sum_i sum_a W_c(a, n, i, r) u_i(a)
// with u_i = v_i minus the atom's streaming velocity (zero without one) and
// vc_i lab-frame (calheatf.F90:368-369, 564, 644, 677-681, 1027-1330)
// each times 1 / (Lx Ly (z2 - z1)), and their total in the reference's order.
// The volume form has only z components (W holds column n = 3); the bulk
// form all three.
//
// Every sum is deterministic: the atoms are cut into a fixed number of
// chunks; one block reduces one (region or class-region row, chunk) pair
// with a fixed-order tree into a partial record, and one thread per
// (region, component) adds the partials in chunk order.  The collision part
// streams the W rows once (HBM-bound: 8 nw bytes per atom per class-region
// row); the transport part re-reads the atom data once per region.
#include "pgk_hf.h"

namespace {

constexpr int kNterm = 1 + PGK_NHF;   // kinetic + potential classes

struct HeatfArgs {
  int natom, nreg, nw, bulk, nsplit, chunk;
  int nsplit_col;             // partial records per collision row
  int nsplit_tr, chunk_tr;    // transport: atom chunks per region and their size
  int ntr_slot;               // transport records per region: nsplit_tr, + 1 with the
                              // pair pass's own transport sums (direct)
  const double *cz;           // the atoms' centre z (wrapped), or null: computed in place
  const double *z1, *z2;
  const int32_t *hftyp, *molid;
  const double *molecom, *pos, *vel, *mass, *molevel;
  const double *vstrm;   // streaming velocities for the kinetic energies, or null
  const double *cvel;    // collision-part velocities: vel - vstrm, or vel
  const void *pe, *w;
  double L[3];
};

// centre-of-mass velocity of each molecule, summed over its atoms in index
// order (calheatf.F90:548-570); molecules are 0-based here
__global__ __launch_bounds__(256) void k_molevel(int nmol, const int32_t *__restrict__ start,
                                                 const int32_t *__restrict__ atoms,
                                                 const double *__restrict__ mass,
                                                 const double *__restrict__ vel,
                                                 double *__restrict__ molevel) {
#pragma clang fp contract(off)
  const int m = blockIdx.x * blockDim.x + threadIdx.x;
  if (m >= nmol) return;
  double mm = 0.0, p0 = 0.0, p1 = 0.0, p2 = 0.0;
  for (int k = start[m]; k < start[m + 1]; k++) {
    const int i = atoms[k];
    mm = mm + mass[i];
    p0 = p0 + mass[i] * vel[3 * i];
    p1 = p1 + mass[i] * vel[3 * i + 1];
    p2 = p2 + mass[i] * vel[3 * i + 2];
  }
  const double inv = 1.0 / mm;
  molevel[3 * m] = p0 * inv;
  molevel[3 * m + 1] = p1 * inv;
  molevel[3 * m + 2] = p2 * inv;
}

// Molecule centres of mass as calcom (calcom.F90:1-147): sum m x over the
// molecule's atoms in index order, times 1/M; a one-atom molecule takes its
// position; then wrapped once into [0, L) per axis.
__global__ __launch_bounds__(256) void k_calcom(int nmol, const int32_t *__restrict__ start,
                                                const int32_t *__restrict__ atoms,
                                                const double *__restrict__ mass,
                                                const double *__restrict__ pos, double3 L,
                                                double *__restrict__ com) {
#pragma clang fp contract(off)
  const int m = blockIdx.x * blockDim.x + threadIdx.x;
  if (m >= nmol) return;
  double c[3];
  if (start[m + 1] - start[m] == 1) {
    const int i = atoms[start[m]];
    c[0] = pos[3 * i];
    c[1] = pos[3 * i + 1];
    c[2] = pos[3 * i + 2];
  } else {
    double mm = 0.0;
    c[0] = c[1] = c[2] = 0.0;
    for (int k = start[m]; k < start[m + 1]; k++) {
      const int i = atoms[k];
      mm = mm + mass[i];
      c[0] = c[0] + mass[i] * pos[3 * i];
      c[1] = c[1] + mass[i] * pos[3 * i + 1];
      c[2] = c[2] + mass[i] * pos[3 * i + 2];
    }
    const double inv = 1.0 / mm;
    c[0] = c[0] * inv;
    c[1] = c[1] * inv;
    c[2] = c[2] * inv;
  }
  const double Ls[3] = {L.x, L.y, L.z};
  for (int d = 0; d < 3; d++) {
    if (c[d] < 0.0) c[d] = c[d] + Ls[d];
    else if (c[d] >= Ls[d]) c[d] = c[d] - Ls[d];
    com[3 * m + d] = c[d];
  }
}

// transport part: block (chunk, region); record [kNterm][3]
template <class AH>
__global__ __launch_bounds__(256) void k_heatf_transport(HeatfArgs a, double *__restrict__ part) {
  const int r = blockIdx.y, split = blockIdx.x;
  const int i0 = split * a.chunk_tr;
  const int i1 = min(a.natom, i0 + a.chunk_tr);
  const double z1 = a.z1[r], z2 = a.z2[r];
  const AH *pe = reinterpret_cast<const AH *>(a.pe);
  double acc[kNterm * 3];
#pragma unroll
  for (int t = 0; t < kNterm * 3; t++) acc[t] = 0.0;
  for (int i = i0 + (int)threadIdx.x; i < i1; i += 256) {
    double cz;
    if (a.cz) {
      cz = a.cz[i];
    } else {
      cz = a.hftyp[i] == PGK_HFTYP_MOLE ? a.molecom[3 * (a.molid[i] - 1) + 2] : a.pos[3 * i + 2];
      if (cz < 0.0) cz = cz + a.L[2];
      else if (cz >= a.L[2]) cz = cz - a.L[2];
    }
    if (!(z1 <= cz && cz <= z2)) continue;
    const double *pv = a.hftyp[i] == PGK_HFTYP_MOLE ? a.molevel + 3 * (a.molid[i] - 1)
                                                    : a.vel + 3 * i;
    // kinetic energy relative to the streaming velocity (calheatf.F90:368-369,
    // 677-681); the transported centre velocity stays lab-frame
    double v[3] = {a.vel[3 * i], a.vel[3 * i + 1], a.vel[3 * i + 2]};
    if (a.vstrm)
      for (int d = 0; d < 3; d++) v[d] = v[d] - a.vstrm[3 * i + d];
    const double ke = 0.5 * a.mass[i] * (v[0] * v[0] + v[1] * v[1] + v[2] * v[2]);
    const double vc[3] = {a.bulk ? pv[0] : 0.0, a.bulk ? pv[1] : 0.0, pv[2]};
#pragma unroll
    for (int c = 0; c < 3; c++) acc[c] += ke * vc[c];
#pragma unroll
    for (int k = 0; k < PGK_NHF; k++) {
      const double e = acc_load(&pe[(size_t)k * a.natom + i]);
#pragma unroll
      for (int c = 0; c < 3; c++) acc[3 * (1 + k) + c] += e * vc[c];
    }
  }
  block_record<kNterm * 3, 4>(acc, part + ((size_t)r * a.ntr_slot + split) * (kNterm * 3));
}

// the atoms' centre z, wrapped into [0, Lz) as the transport part tests it
// (ATOM: the atom, MOLE: its molecule's centre of mass), once per pass
// instead of once per region
__global__ __launch_bounds__(256) void k_centre_z(HeatfArgs a, double *__restrict__ cz) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= a.natom) return;
  double z = a.hftyp[i] == PGK_HFTYP_MOLE ? a.molecom[3 * (a.molid[i] - 1) + 2] : a.pos[3 * i + 2];
  if (z < 0.0) z = z + a.L[2];
  else if (z >= a.L[2]) z = z - a.L[2];
  cz[i] = z;
}

// collision part: block (chunk, class-region row); record [3]
template <class AH>
__global__ __launch_bounds__(256) void k_heatf_collision(HeatfArgs a, double *__restrict__ part) {
  const int row = blockIdx.y, split = blockIdx.x;
  const int i0 = split * a.chunk;
  const int i1 = min(a.natom, i0 + a.chunk);
  const AH *W = reinterpret_cast<const AH *>(a.w) + (size_t)row * a.natom * a.nw;
  double acc[3] = {0.0, 0.0, 0.0};
  for (int i = i0 + (int)threadIdx.x; i < i1; i += 256) {
    const double v0 = a.cvel[3 * i], v1 = a.cvel[3 * i + 1], v2 = a.cvel[3 * i + 2];
    const AH *w = W + (size_t)i * a.nw;
    if (a.nw == 9) {
#pragma unroll
      for (int n = 0; n < 3; n++)
        acc[n] += acc_load(&w[3 * n]) * v0 + acc_load(&w[3 * n + 1]) * v1 +
                  acc_load(&w[3 * n + 2]) * v2;
    } else {
      acc[2] += acc_load(&w[0]) * v0 + acc_load(&w[1]) * v1 + acc_load(&w[2]) * v2;
    }
  }
  block_record<3, 4>(acc, part + ((size_t)row * a.nsplit + split) * 3);
}

// partials in chunk order, the volume factor, the total (calheatf.F90:1334-1436).
// One block per region; lane l < 51 of each wave owns (term l / 3, component
// l % 3), wave w sums the chunks k = w, w + 4, ... and wave 0 adds the four
// partial sums in order, so the result does not depend on timing.
__global__ __launch_bounds__(256) void k_heatf_final(HeatfArgs a, const double *__restrict__ ptr,
                                                     const double *__restrict__ pcol,
                                                     double *__restrict__ out) {
#pragma clang fp contract(off)
  constexpr int kTc = PGK_HFJ_TOT * 3;   // 51 (term, component) pairs
  __shared__ double part[4][kTc];
  __shared__ double fin[kTc];
  const int r = blockIdx.x, wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  if (lane < kTc) {
    const int t = lane / 3, c = lane % 3;
    double s = 0.0;
    if (t < kNterm) {
      for (int k = wave; k < a.ntr_slot; k += 4)
        s += ptr[((size_t)r * a.ntr_slot + k) * (kNterm * 3) + 3 * t + c];
    } else {
      const size_t row = (size_t)(t - kNterm) * a.nreg + r;
      for (int k = wave; k < a.nsplit_col; k += 4) s += pcol[(row * a.nsplit_col + k) * 3 + c];
    }
    part[wave][lane] = s;
  }
  __syncthreads();
  const double vol_inv = ((1.0 / a.L[0]) * (1.0 / a.L[1])) / (a.z2[r] - a.z1[r]);
  double *o = out + (size_t)r * PGK_NHFJ * 3;
  if (threadIdx.x < kTc) {
    const double s = (((part[0][threadIdx.x] + part[1][threadIdx.x]) + part[2][threadIdx.x]) +
                      part[3][threadIdx.x]) * vol_inv;
    fin[threadIdx.x] = s;
    o[threadIdx.x] = s;   // term t, component c at 3 t + c
  }
  __syncthreads();
  if (threadIdx.x < 3) {
    double tot = fin[threadIdx.x];
    for (int t = 1; t < PGK_HFJ_TOT; t++) tot = tot + fin[3 * t + threadIdx.x];
    o[3 * PGK_HFJ_TOT + threadIdx.x] = tot;
  }
}

// flux form: collision row (class, region) = the bonded terms' sums in jcol
// (heat-flux accumulator type AH) plus, for the pair classes, the pair pass's
// per-block records in block order (block (row, component) per launch block,
// fixed-order tree).  direct (bulk, one region): the pair pass's transport
// sums too, into the transport record slot after the atom chunks' (tslot),
// zero for the other terms.
template <class AH>
__global__ __launch_bounds__(256) void k_heatf_collect(int nreg, int bulk, int nblk,
                                                       const AH *__restrict__ jcol,
                                                       const double *__restrict__ jpart,
                                                       double *__restrict__ pcol, int direct,
                                                       double *__restrict__ tslot) {
  const int row = blockIdx.x / 3, c = blockIdx.x % 3;   // row = class * nreg + region
  const int k = row / nreg, r = row % nreg;
  const int q = k - PGK_HF_ELIN;
  const bool pair = q >= 0 && q < 4;
  double v[2] = {0.0, 0.0};
  if (pair && (bulk || c == 2)) {
    const int stride = bulk ? 24 : 4 * nreg;
    const int off = bulk ? 3 * q + c : q * nreg + r;
    for (int b = threadIdx.x; b < nblk; b += 256) {
      v[0] += jpart[(size_t)b * stride + off];
      if (direct) v[1] += jpart[(size_t)b * stride + 12 + off];
    }
  }
  __shared__ double rec[PGK_REC];
  block_record<2, 4>(v, rec);
  __syncthreads();
  if (threadIdx.x == 0) {
    pcol[(size_t)row * 3 + c] = acc_load(&jcol[(size_t)row * 3 + c]) + rec[0];
    if (direct) {   // term 1 + k of the transport record (term 0: kinetic)
      tslot[3 * (1 + k) + c] = pair ? rec[1] : 0.0;
      if (k == 0) tslot[c] = 0.0;
    }
  }
}

// clears up to three 8-byte arrays in one launch (the heat-flux sums a pass
// adds to), in place of one fill per array
__global__ __launch_bounds__(256) void k_zero3(uint64_t *__restrict__ p0, size_t n0,
                                               uint64_t *__restrict__ p1, size_t n1,
                                               uint64_t *__restrict__ p2, size_t n2) {
  const size_t stride = (size_t)gridDim.x * blockDim.x;
  for (size_t k = (size_t)blockIdx.x * blockDim.x + threadIdx.x; k < n0 + n1 + n2; k += stride) {
    if (k < n0) p0[k] = 0;
    else if (k < n0 + n1) p1[k - n0] = 0;
    else p2[k - n0 - n1] = 0;
  }
}

// velocities relative to the streaming velocities
__global__ __launch_bounds__(256) void k_peculiar(int n3, const double *__restrict__ vel,
                                                  const double *__restrict__ vstrm,
                                                  double *__restrict__ out) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k < n3) out[k] = vel[k] - vstrm[k];
}

template <int PREC>
int launch_heatf(pgk_ctx *c, const HeatfArgs &a, double *ptr, double *pcol) {
  using AH = typename HfAcc<PREC>::type;
  hipStream_t st = c->stream;
  k_heatf_transport<AH><<<dim3(a.nsplit_tr, a.nreg), 256, 0, st>>>(a, ptr);
  HIP_CHECK(hipGetLastError());
  if (c->hfdev.flux) {
    const int direct = a.ntr_slot > a.nsplit_tr;   // one region: its slot after the chunks'
    k_heatf_collect<AH><<<PGK_NHF * a.nreg * 3, 256, 0, st>>>(
        a.nreg, a.bulk, pgk_grid_blocks(pgk_pair_waves(c), 4),
        reinterpret_cast<const AH *>(c->hfdev.jcol), c->hfdev.jpart, pcol, direct,
        ptr + (size_t)a.nsplit_tr * (kNterm * 3));
  } else {
    k_heatf_collision<AH><<<dim3(a.nsplit, a.nreg * PGK_NHF), 256, 0, st>>>(a, pcol);
  }
  HIP_CHECK(hipGetLastError());
  return PGK_OK;
}

}  // namespace

int pgk_zero3_launch(hipStream_t st, void *p0, size_t n0, void *p1, size_t n1, void *p2,
                     size_t n2) {
  const size_t n = n0 + n1 + n2;
  if (n == 0) return PGK_OK;
  const int nb = (int)std::min<size_t>((n + 255) / 256, 2048);
  k_zero3<<<nb, 256, 0, st>>>((uint64_t *)p0, n0, (uint64_t *)p1, n1, (uint64_t *)p2, n2);
  HIP_CHECK(hipGetLastError());
  return PGK_OK;
}

// The velocities of the collision part: d_vel, or (with streaming
// velocities) d_vel - d_vstrm queued into the context's buffer.
int pgk_heatf_collision_vel(pgk_ctx *c, const double *d_vel, const double *d_vstrm,
                            const double **out) {
  if (!d_vstrm) {
    *out = d_vel;
    return PGK_OK;
  }
  const int n3 = 3 * c->natom;
  int rc;
  if ((rc = grow_dev((void **)&c->d_hj_vpec, &c->hj_vpec_cap, sizeof(double) * n3))) return rc;
  k_peculiar<<<pgk_grid_blocks(n3, 256), 256, 0, c->stream>>>(n3, d_vel, d_vstrm, c->d_hj_vpec);
  HIP_CHECK(hipGetLastError());
  *out = c->d_hj_vpec;
  return PGK_OK;
}

// Queue calheatf over the context's heat-flux arrays; velocities and masses
// are already in d_vel / d_mass (device, natom entries), the result goes to
// c->d_hj_out.
int pgk_heatf_launch(pgk_ctx *c, const double *d_vel, const double *d_mass,
                     const double *d_vstrm, double *out) {
  const HfDev &h = c->hfdev;
  const int n = c->natom;
  HeatfArgs a;
  a.natom = n;
  a.nreg = h.nreg;
  a.nw = h.nw;
  a.bulk = h.mode == PGK_HF_BULK;
  // chunks: about 2048 collision blocks, at least 256 atoms each
  const int rows = PGK_NHF * h.nreg;
  int nsplit = (2048 + rows - 1) / rows;
  nsplit = std::max(1, std::min(nsplit, (n + 255) / 256));
  a.nsplit = nsplit;
  a.chunk = (n + nsplit - 1) / nsplit;
  a.nsplit_col = h.flux ? 1 : nsplit;
  // transport: every region's blocks scan the atoms, about 2048 blocks in all
  int nsplit_tr = (2048 + h.nreg - 1) / h.nreg;
  nsplit_tr = std::max(1, std::min(nsplit_tr, (n + 255) / 256));
  a.nsplit_tr = nsplit_tr;
  a.chunk_tr = (n + nsplit_tr - 1) / nsplit_tr;
  a.ntr_slot = nsplit_tr + ((h.flux && h.direct_tr) ? 1 : 0);
  a.cz = nullptr;
  a.z1 = h.z1;
  a.z2 = h.z2;
  a.hftyp = h.hftyp;
  a.molid = h.molid;
  a.molecom = h.molecom;
  a.pos = h.pos;
  a.vel = d_vel;
  a.vstrm = d_vstrm;
  a.cvel = d_vel;
  a.mass = d_mass;
  a.molevel = c->d_hj_molevel;
  a.pe = h.pe;
  a.w = h.w;
  for (int d = 0; d < 3; d++) a.L[d] = h.L[d];
  const size_t ntr = (size_t)h.nreg * a.ntr_slot * kNterm * 3;
  const size_t ncol = (size_t)rows * a.nsplit_col * 3;
  int rc;
  if (!h.flux && (rc = pgk_heatf_collision_vel(c, d_vel, d_vstrm, &a.cvel))) return rc;
  if ((rc = grow_dev((void **)&c->d_hj_part, &c->hj_part_cap, sizeof(double) * (ntr + ncol))) ||
      (rc = grow_dev((void **)&c->d_hj_out, &c->hj_out_cap,
                     sizeof(double) * (size_t)h.nreg * PGK_NHFJ * 3)))
    return rc;
  if (h.molecom) {
    k_molevel<<<pgk_grid_blocks(c->hj_nmol, 256), 256, 0, c->stream>>>(
        c->hj_nmol, c->d_hj_mol_start, c->d_hj_mol_atoms, d_mass, d_vel, c->d_hj_molevel);
    HIP_CHECK(hipGetLastError());
  }
  if (h.nreg > 1) {   // many regions: the centres once, not once per region
    if ((rc = grow_dev((void **)&c->d_hj_cz, &c->hj_cz_cap, sizeof(double) * n))) return rc;
    k_centre_z<<<pgk_grid_blocks(n, 256), 256, 0, c->stream>>>(a, c->d_hj_cz);
    HIP_CHECK(hipGetLastError());
    a.cz = c->d_hj_cz;
  }
  double *ptr = c->d_hj_part, *pcol = c->d_hj_part + ntr;
  switch (c->precision) {
    case PGK_PREC_MIXED: rc = launch_heatf<PGK_PREC_MIXED>(c, a, ptr, pcol); break;
    case PGK_PREC_FP64: rc = launch_heatf<PGK_PREC_FP64>(c, a, ptr, pcol); break;
    default: rc = launch_heatf<PGK_PREC_FP32>(c, a, ptr, pcol); break;
  }
  if (rc) return rc;
  k_heatf_final<<<h.nreg, 256, 0, c->stream>>>(a, ptr, pcol, out ? out : c->d_hj_out);
  HIP_CHECK(hipGetLastError());
  return PGK_OK;
}

int pgk_calcom_launch(pgk_ctx *c, const double *d_mass, double *d_com) {
  if (c->hj_nmol <= 0) return PGK_OK;
  k_calcom<<<pgk_grid_blocks(c->hj_nmol, 256), 256, 0, c->stream>>>(
      c->hj_nmol, c->d_hj_mol_start, c->d_hj_mol_atoms, d_mass, c->d_pos,
      make_double3(c->box[0], c->box[1], c->box[2]), d_com);
  HIP_CHECK(hipGetLastError());
  return PGK_OK;
}
