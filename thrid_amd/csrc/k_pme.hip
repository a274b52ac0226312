// k_pme.hip - smooth particle-mesh Ewald reciprocal part.
//
// Restates fft_pmep (fft_pme.F90:155-270):
//   get_scaled_fractionals (:273-313)  fr_d = K_d (w - floor w), w = x_d/L_d
//   fill_bspline (fft_pme_init.F90:216-260) at fr - int(fr)
//   fill_charge_grid (:359-422)        Q(int(fr)-p+1 .. int(fr)) += q th1 th2 th3
//   fft_back  (pubz3d isign -1)        forward transform (e^{-2 pi i jk/n})
//   scalar_sum (:544-690)              eterm = exp(-pi^2 m^2/a^2)/(pi V B m^2),
//                                      zero outside |m| <= rkcut and at m = 0
//                                      (the m = 0 value is left untouched),
//                                      E = 1/2 sum eterm |S|^2, virial, Q *= eterm
//   fft_forward (pubz3d isign +1)      unnormalised backward transform
//   grad_sum (:872-937)                F = -q sum Q (K1 dth1 th2 th3, ...) / L
//
// MI355X form: the charge grid is real, so the x transform is real-to-complex
// (two grid rows packed into one complex line) and only the half spectrum
// h = 0..K1/2 is kept.  Modes off the h = 0 and h = K1/2 planes stand for
// themselves and their mirror image; their energy counts twice and each
// off-diagonal virial product counts (1 + s_a s_b) times, where s_d = -1
// unless index d sits on its Nyquist plane (whose mirror image keeps the
// reference's negative representative m_d = -K_d/2).  The y and z transforms
// are batched Stockham radix-2/3/4/5 FFTs in LDS; the convolution is fused
// between the forward and backward z transforms, so the spectrum makes one
// round trip through HBM.
#include <hipcub/hipcub.hpp>
#include <algorithm>
#include <type_traits>

#include "pgk_device.h"

namespace {

template <class R> struct C2T;
template <> struct C2T<double> { using type = double2; };
template <> struct C2T<float> { using type = float2; };

template <class C>
__device__ __forceinline__ C cmul(C a, C b) {
  C r;
  r.x = a.x * b.x - a.y * b.y;
  r.y = a.x * b.y + a.y * b.x;
  return r;
}
template <class C>
__device__ __forceinline__ C cadd(C a, C b) {
  C r;
  r.x = a.x + b.x;
  r.y = a.y + b.y;
  return r;
}
template <class C>
__device__ __forceinline__ C csub(C a, C b) {
  C r;
  r.x = a.x - b.x;
  r.y = a.y - b.y;
  return r;
}
// multiply by -i (forward) or +i (inverse): sgn = -1 forward
template <class C>
__device__ __forceinline__ C cmuli(C a, int sgn) {
  C r;
  if (sgn < 0) {
    r.x = a.y;
    r.y = -a.x;
  } else {
    r.x = -a.y;
    r.y = a.x;
  }
  return r;
}

// in-register DFT of size R with sign sgn (-1 forward, +1 inverse)
template <class C>
__device__ __forceinline__ void dft(C *v, int R, int sgn) {
  using T = decltype(v[0].x);
  if (R == 2) {
    C a = v[0], b = v[1];
    v[0] = cadd(a, b);
    v[1] = csub(a, b);
  } else if (R == 4) {
    C a0 = cadd(v[0], v[2]), a1 = csub(v[0], v[2]);
    C b0 = cadd(v[1], v[3]), b1 = cmuli(csub(v[1], v[3]), sgn);
    v[0] = cadd(a0, b0);
    v[2] = csub(a0, b0);
    v[1] = cadd(a1, b1);
    v[3] = csub(a1, b1);
  } else if (R == 3) {
    const T c1 = T(-0.5), s1 = T(0.86602540378443864676) * T(sgn);
    C t1 = cadd(v[1], v[2]), t2 = csub(v[1], v[2]);
    C m1;
    m1.x = v[0].x + c1 * t1.x;
    m1.y = v[0].y + c1 * t1.y;
    C m2;
    m2.x = -s1 * t2.y;
    m2.y = s1 * t2.x;
    v[0] = cadd(v[0], t1);
    v[1] = cadd(m1, m2);
    v[2] = csub(m1, m2);
  } else if (R == 5) {
    const T c1 = T(0.30901699437494742410), c2 = T(-0.80901699437494742410);
    const T s1 = T(0.95105651629515357212) * T(sgn), s2 = T(0.58778525229247312917) * T(sgn);
    C t1 = cadd(v[1], v[4]), t2 = cadd(v[2], v[3]);
    C t3 = csub(v[1], v[4]), t4 = csub(v[2], v[3]);
    C a1, a2, b1, b2;
    a1.x = v[0].x + c1 * t1.x + c2 * t2.x;
    a1.y = v[0].y + c1 * t1.y + c2 * t2.y;
    a2.x = v[0].x + c2 * t1.x + c1 * t2.x;
    a2.y = v[0].y + c2 * t1.y + c1 * t2.y;
    // b = i*(s1 t3 + s2 t4) and i*(s2 t3 - s1 t4)
    b1.x = -(s1 * t3.y + s2 * t4.y);
    b1.y = s1 * t3.x + s2 * t4.x;
    b2.x = -(s2 * t3.y - s1 * t4.y);
    b2.y = s2 * t3.x - s1 * t4.x;
    v[0] = cadd(v[0], cadd(t1, t2));
    v[1] = cadd(a1, b1);
    v[4] = csub(a1, b1);
    v[2] = cadd(a2, b2);
    v[3] = csub(a2, b2);
  }
}

struct FftPlan {
  int n;
  int nrad;
  int rad[24];
};

// q = t / d, r = t % d for 0 <= t < 2^22 and 1 <= d < 2^16 without integer
// division: a float reciprocal estimate and one correction step
__device__ __forceinline__ int fdiv(int t, int d, float rd) {
  int q = (int)((float)t * rd);
  int r = t - q * d;
  q += (r >= d) - (r < 0);
  return q;
}

// One Stockham pass of radix R over nl lines (compile-time R keeps the
// butterfly operands in registers).
template <int R, class C>
__device__ __forceinline__ void stockham_pass(const C *pa, C *pb, int nl, int n, int ls, int Ns,
                                              const C *tw, int sgn) {
  const int nb = n / R;
  const float rnb = 1.0f / (float)nb, rns = 1.0f / (float)Ns;
  const int tstride = n / (Ns * R);
  for (int t = threadIdx.x; t < nl * nb; t += blockDim.x) {
    const int line = fdiv(t, nb, rnb), j = t - line * nb;
    const C *src = pa + line * ls;
    C *dst = pb + line * ls;
    const int jq = fdiv(j, Ns, rns), k = j - jq * Ns;
    C v[R];
    v[0] = src[j];
    const int tk = k * tstride;   // (q k tstride) < n: no wrap needed
#pragma unroll
    for (int q = 1; q < R; q++) {
      C w = tw[q * tk];
      if (sgn > 0) w.y = -w.y;
      v[q] = cmul(src[j + q * nb], w);
    }
    dft(v, R, sgn);
    const int idxD = jq * Ns * R + k;
#pragma unroll
    for (int q = 0; q < R; q++) dst[idxD + q * Ns] = v[q];
  }
}

// Batched Stockham FFT of nl lines of length p.n held in LDS with line
// stride ls (n + 1 where lines are filled column-wise, so that lanes walking
// across lines hit different banks).  Result ends in *pa (the buffers are
// swapped pass by pass).
template <class C>
__device__ void fft_lds(C *&pa, C *&pb, int nl, int ls, const FftPlan &p, const C *tw, int sgn) {
  const int n = p.n;
  int Ns = 1;
  for (int s = 0; s < p.nrad; s++) {
    const int R = p.rad[s];
    switch (R) {
      case 4: stockham_pass<4>(pa, pb, nl, n, ls, Ns, tw, sgn); break;
      case 2: stockham_pass<2>(pa, pb, nl, n, ls, Ns, tw, sgn); break;
      case 3: stockham_pass<3>(pa, pb, nl, n, ls, Ns, tw, sgn); break;
      default: stockham_pass<5>(pa, pb, nl, n, ls, Ns, tw, sgn); break;
    }
    __syncthreads();
    C *t = pa;
    pa = pb;
    pb = t;
    Ns *= R;
  }
}


// ---- register FFT for line lengths N = N1 * N2 in {32, 64, 128, 256} ----
// N2 lanes share a line; lane t starts with elements t + N2 j (j < N1),
// transforms them with an in-register N1-point DFT, applies the twiddles
// W_N^(t k1), swaps through LDS and finishes with N1/N2 in-register N2-point
// DFTs.  It ends holding elements (t + N2 q) + N1 k2.  One wavefront holds
// 64/N2 lines and needs no barrier wider than itself.

// cos and sin of 2 pi k / 16, k < 8
__host__ __device__ constexpr double cos16(int k) {
  return k == 0 ? 1.0
       : k == 1 ? 0.92387953251128675613
       : k == 2 ? 0.70710678118654752440
       : k == 3 ? 0.38268343236508977173
       : k == 4 ? 0.0
       : k == 5 ? -0.38268343236508977173
       : k == 6 ? -0.70710678118654752440
                : -0.92387953251128675613;
}
__host__ __device__ constexpr double sin16(int k) { return k <= 4 ? cos16(4 - k) : cos16(k - 4); }

// a * exp(SGN 2 pi i k / M), M <= 16
template <int M, int SGN, class C>
__device__ __forceinline__ C twc(C a, int k) {
  using T = decltype(a.x);
  const int idx = k * (16 / M);
  if (idx == 0) return a;
  if (idx == 4) return cmuli(a, SGN);
  const T c = (T)cos16(idx), sn = (T)(SGN * sin16(idx));
  C r;
  r.x = a.x * c - a.y * sn;
  r.y = a.x * sn + a.y * c;
  return r;
}

// in-register M-point DFT (radix-2, decimation in time), natural order
template <int M, int SGN, class C>
__device__ __forceinline__ void dft_r(C *v) {
  if constexpr (M > 1) {
    C e[M / 2], o[M / 2];
#pragma unroll
    for (int i = 0; i < M / 2; i++) {
      e[i] = v[2 * i];
      o[i] = v[2 * i + 1];
    }
    dft_r<M / 2, SGN>(e);
    dft_r<M / 2, SGN>(o);
#pragma unroll
    for (int k = 0; k < M / 2; k++) {
      const C tk = twc<M, SGN>(o[k], k);
      v[k] = cadd(e[k], tk);
      v[k + M / 2] = csub(e[k], tk);
    }
  }
}

template <int N1_, int N2_>
struct RegFft {
  static constexpr int N1 = N1_, N2 = N2_, N = N1_ * N2_;
  static constexpr int LPW = 64 / N2_;          // lines per wavefront
  static constexpr int XS = 65;                 // exchange row stride (complex)
  static constexpr int NS = N + N2_;            // natural-order line stride (complex)
  static constexpr int BUF = (N1_ * XS > LPW * NS) ? N1_ * XS : LPW * NS;
};

// tw: W_N^m = exp(-2 pi i m / N), m < N, in LDS; ex: the wave's exchange
// buffer (RegFft::BUF complex)
template <class F, int SGN, class C>
__device__ __forceinline__ void fft_reg(C (&x)[F::N1], C *ex, const C *tw, int t, int l) {
  dft_r<F::N1, SGN>(x);
#pragma unroll
  for (int k1 = 1; k1 < F::N1; k1++) {
    C w = tw[t * k1];
    if (SGN > 0) w.y = -w.y;
    x[k1] = cmul(x[k1], w);
  }
#pragma unroll
  for (int k1 = 0; k1 < F::N1; k1++) ex[k1 * F::XS + l * F::N2 + t] = x[k1];
  __syncthreads();
#pragma unroll
  for (int q = 0; q < F::N1 / F::N2; q++) {
    C v[F::N2];
#pragma unroll
    for (int t2 = 0; t2 < F::N2; t2++) v[t2] = ex[(t + F::N2 * q) * F::XS + l * F::N2 + t2];
    dft_r<F::N2, SGN>(v);
#pragma unroll
    for (int k2 = 0; k2 < F::N2; k2++) x[q * F::N2 + k2] = v[k2];
  }
  __syncthreads();
}

// B-spline weights of order P at w (fill_bspline, fft_pme_init.F90:216-260)
template <class R, int P>
__device__ __forceinline__ void bspline_t(R w, R *A, R *D) {
  // A[1..P], D[1..P]
  A[P] = R(0);
  A[2] = w;
  A[1] = R(1) - w;
#pragma unroll
  for (int k = 3; k <= P - 1; k++) {
    const R div = R(1) / R(k - 1);
    A[k] = div * w * A[k - 1];
#pragma unroll
    for (int j = 1; j <= k - 2; j++)
      A[k - j] = div * ((w + R(j)) * A[k - j - 1] + (R(k - j) - w) * A[k - j]);
    A[1] = div * (R(1) - w) * A[1];
  }
  D[1] = -A[1];
#pragma unroll
  for (int j = 2; j <= P; j++) D[j] = A[j - 1] - A[j];
  const R div = R(1) / R(P - 1);
  A[P] = div * w * A[P - 1];
#pragma unroll
  for (int j = 1; j <= P - 2; j++)
    A[P - j] = div * ((w + R(j)) * A[P - j - 1] + (R(P - j) - w) * A[P - j]);
  A[1] = div * (R(1) - w) * A[1];
}

__device__ __forceinline__ double msq_nc(double a, double b, double c) {
#pragma clang fp contract(off)
  return a * a + b * b + c * c;
}

struct PmeGeom {
  int K[3];
  int H;            // K1/2 + 1
  double recip[3];
  int order;
};

#define PGK_MAXSRC 125   // source bins per tile (5 per axis)
#define PGK_SPREAD_MAXVIS 512   // kept atom visits per tile listed in LDS

struct TileGeom {
  int T[3], nt[3], nsrc[3];   // tile size, tiles per axis, source bins per axis
  int ntile;
};

__device__ __forceinline__ int tile_home(const int4 &b, const PmeGeom &g, const TileGeom &tg) {
  const int bx = b.x < 0 ? b.x + g.K[0] : b.x;
  const int by = b.y < 0 ? b.y + g.K[1] : b.y;
  const int bz = b.z < 0 ? b.z + g.K[2] : b.z;
  return ((bz / tg.T[2]) * tg.nt[1] + by / tg.T[1]) * tg.nt[0] + bx / tg.T[0];
}


template <class R, int P>
__device__ __forceinline__ void atom_frac_t(const double *pos, int i, const PmeGeom &g, int *i0,
                                            R (*th)[P + 1], R (*dth)[P + 1]) {
#pragma unroll
  for (int d = 0; d < 3; d++) {
    double w = pos[3 * i + d] * g.recip[d];
    w = w - floor(w);
    double fr = (double)g.K[d] * w;
    if (!(fr >= 0.0 && fr < (double)g.K[d])) fr = (fr >= (double)g.K[d]) ? (double)g.K[d] - 1.0 : 0.0;
    int ifr = (int)fr;
    R ww = (R)(fr - (double)ifr);
    bspline_t<R, P>(ww, th[d], dth[d]);
    i0[d] = ifr - P;
  }
}

// Per-slot preparation ahead of the force kernels, in the list's spatial order
// (order[k] = atom, or -1 for padding; identity without a list):
//   - B-spline weights th[(d*P + q)*nitem + k] and the first stencil index
//     base[k] = int(fr) - P + 1 per axis (get_bspline_coeffs and
//     fill_charge_grid, fft_pme.F90:316-422), base.w = atom;
//   - the slot's rank in its home tile bin: the block counts in LDS (open
//     addressing on the tile id) and takes one global atomic per tile;
//   - with PACK, the pair kernel's (x, y, z, q) and 0-based LJ type per slot.
template <int PREC, int P, bool PACK>
__global__ __launch_bounds__(256) void k_prep(int nitem, const int32_t *__restrict__ order,
                                              const double *__restrict__ pos,
                                              const double *__restrict__ q,
                                              const int32_t *__restrict__ type, PmeGeom g,
                                              TileGeom tg,
                                              typename Prec<PREC>::real *__restrict__ th,
                                              int4 *__restrict__ base,
                                              int32_t *__restrict__ count,
                                              int32_t *__restrict__ rank,
                                              typename Prec<PREC>::real4 *__restrict__ xq,
                                              int32_t *__restrict__ ct) {
  using R = typename Prec<PREC>::real;
  using R4 = typename Prec<PREC>::real4;
  __shared__ int s_key[512], s_cnt[512], s_base[512];
  for (int i = threadIdx.x; i < 512; i += blockDim.x) {
    s_key[i] = -1;
    s_cnt[i] = 0;
  }
  __syncthreads();
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  int t = -1;
  if (k < nitem) {
    const int a = order ? order[k] : k;
    int4 b;
    b.w = a;
    if (a < 0) {
      b.x = b.y = b.z = 0;
      if (PACK) {
        R4 z;
        z.x = z.y = z.z = z.w = R(0);
        xq[k] = z;
        ct[k] = 0;
      }
    } else {
      R tt[3][P + 1], dt[3][P + 1];
      int i0[3];
      atom_frac_t<R, P>(pos, a, g, i0, tt, dt);
#pragma unroll
      for (int d = 0; d < 3; d++)
#pragma unroll
        for (int qq = 0; qq < P; qq++) th[(size_t)(d * P + qq) * nitem + k] = tt[d][qq + 1];
      b.x = i0[0] + 1;
      b.y = i0[1] + 1;
      b.z = i0[2] + 1;
      t = tile_home(b, g, tg);
      if (PACK) {
        R4 v;
        v.x = (R)pos[3 * a];
        v.y = (R)pos[3 * a + 1];
        v.z = (R)pos[3 * a + 2];
        v.w = (R)q[a];
        xq[k] = v;
        ct[k] = type[a] - 1;   // 0-based LJ type index
      }
    }
    base[k] = b;
  }
  int slot = -1, r = -1;
  if (t >= 0) {
    int h = (int)(((unsigned)t * 2654435761u) >> 23);   // 9-bit hash
    for (int probe = 0; probe < 512; probe++) {
      const int old = atomicCAS(&s_key[h], -1, t);
      if (old == -1 || old == t) {
        slot = h;
        break;
      }
      h = (h + 1) & 511;
    }
    r = atomicAdd(&s_cnt[slot], 1);
  }
  __syncthreads();
  for (int i = threadIdx.x; i < 512; i += blockDim.x)
    if (s_key[i] >= 0) s_base[i] = atomicAdd(&count[s_key[i]], s_cnt[i]);
  __syncthreads();
  if (k < nitem) rank[k] = (t >= 0) ? s_base[slot] + r : -1;
}

// Forces from the potential grid (grad_sum, fft_pme.F90:515-590): four
// lanes per slot, lane s taking the stencil planes a3 = s, s + 4, ... (four
// times the threads of one lane per atom, so the grid loads of more atoms are
// in flight at once); the lanes' partial forces are added in a fixed tree and
// the first lane of the four adds the force to the atom's accumulator.
template <int PREC, bool SPLIT, int P>
__global__ __launch_bounds__(256) void k_gather_t(int nitem, const int32_t *__restrict__ order,
                                                  const double *__restrict__ pos,
                                                  const double *__restrict__ q, PmeGeom g,
                                                  const void *gridv,
                                                  typename Prec<PREC>::acc *fat,
                                                  typename Prec<PREC>::acc *fat_c) {
  using R = typename Prec<PREC>::real;
  const int t = xcd_block() * blockDim.x + threadIdx.x;
  const int k = t >> 2, sub = t & 3;
  // whole groups of four lanes leave together, so the shuffles below see
  // their partners
  if (k >= nitem) return;
  const int i = order ? order[k] : k;
  if (i < 0) return;
  const R *grid = reinterpret_cast<const R *>(gridv);
  R th[3][P + 1], dth[3][P + 1];
  int i0[3];
  atom_frac_t<R, P>(pos, i, g, i0, th, dth);
  R f1 = 0, f2 = 0, f3 = 0;
  const R qi = (R)q[i];
  const R K1 = (R)g.K[0], K2 = (R)g.K[1], K3 = (R)g.K[2];
  int ix[P];
#pragma unroll
  for (int a1 = 1; a1 <= P; a1++) {
    int ii = i0[0] + a1;
    ix[a1 - 1] = (ii < 0) ? ii + g.K[0] : ii;
  }
  for (int a3 = 1 + sub; a3 <= P; a3 += 4) {
    int kk = i0[2] + a3;
    if (kk < 0) kk += g.K[2];
    R t3 = 0, d3 = 0;
#pragma unroll
    for (int a = 1; a <= P; a++)
      if (a == a3) {
        t3 = th[2][a];
        d3 = dth[2][a];
      }
#pragma unroll
    for (int a2 = 1; a2 <= P; a2++) {
      int jj = i0[1] + a2;
      if (jj < 0) jj += g.K[1];
      const R *row = grid + ((size_t)kk * g.K[1] + jj) * g.K[0];
      const R t23 = th[1][a2] * t3;
      const R d23 = dth[1][a2] * t3;
      const R t2d3 = th[1][a2] * d3;
#pragma unroll
      for (int a1 = 1; a1 <= P; a1++) {
        const R term = qi * row[ix[a1 - 1]];
        f1 = f1 - K1 * term * dth[0][a1] * t23;
        f2 = f2 - K2 * term * th[0][a1] * d23;
        f3 = f3 - K3 * term * th[0][a1] * t2d3;
      }
    }
  }
  f1 += __shfl_xor(f1, 1, 64);
  f2 += __shfl_xor(f2, 1, 64);
  f3 += __shfl_xor(f3, 1, 64);
  f1 += __shfl_xor(f1, 2, 64);
  f2 += __shfl_xor(f2, 2, 64);
  f3 += __shfl_xor(f3, 2, 64);
  if (sub != 0) return;
  double fx = g.recip[0] * (double)f1, fy = g.recip[1] * (double)f2, fz = g.recip[2] * (double)f3;
  acc_add(&fat[3 * i + 0], fx);
  acc_add(&fat[3 * i + 1], fy);
  acc_add(&fat[3 * i + 2], fz);
  if (SPLIT) {
    acc_add(&fat_c[3 * i + 0], fx);
    acc_add(&fat_c[3 * i + 1], fy);
    acc_add(&fat_c[3 * i + 2], fz);
  }
}

// Per-atom reciprocal energy from the potential grid the last call left in
// d_pgrid: e_i = q_i phi(r_i) / 2 with phi(r_i) = sum theta1 theta2 theta3
// grid, the same interpolation as k_gather_t's with the B-splines in place
// of their derivatives (F_i = -q_i grad phi(r_i)), so sum_i e_i is the
// reciprocal energy.  One thread per atom, atom order.
template <int PREC, int P>
__global__ __launch_bounds__(256) void k_phi_atoms(int n, const double *__restrict__ pos,
                                                   const double *__restrict__ q, PmeGeom g,
                                                   const void *gridv, double *__restrict__ e) {
  using R = typename Prec<PREC>::real;
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const R *grid = reinterpret_cast<const R *>(gridv);
  R th[3][P + 1], dth[3][P + 1];
  int i0[3];
  atom_frac_t<R, P>(pos, i, g, i0, th, dth);
  R phi = 0;
#pragma unroll
  for (int a3 = 1; a3 <= P; a3++) {
    int kk = i0[2] + a3;
    if (kk < 0) kk += g.K[2];
#pragma unroll
    for (int a2 = 1; a2 <= P; a2++) {
      int jj = i0[1] + a2;
      if (jj < 0) jj += g.K[1];
      const R *row = grid + ((size_t)kk * g.K[1] + jj) * g.K[0];
      const R t23 = th[1][a2] * th[2][a3];
#pragma unroll
      for (int a1 = 1; a1 <= P; a1++) {
        int ii = i0[0] + a1;
        if (ii < 0) ii += g.K[0];
        phi += row[ii] * th[0][a1] * t23;
      }
    }
  }
  e[i] = 0.5 * q[i] * (double)phi;
}

// ---- charge spreading (fill_charge_grid, fft_pme.F90:359-422) -------------
// Owner-computes: the grid is cut into equal tiles of T0 x T1 x T2 points
// (T = largest divisor of K that is <= 8).  Atoms are binned by the tile of
// their first stencil point; a stencil of p points reaches nb = (T+p-2)/T
// further tiles, so tile t reads the bins t, t-1, ..., t-nb of each axis.
// One workgroup per tile accumulates the points it owns in LDS (int64 fixed
// point in the mixed form: order-independent) and writes each grid point
// once, in the compute type.  No device atomics, no clearing pass.
// exclusive scan of count[0..n) into start[0..n] (start[n] = total), one
// block: each thread sums a contiguous chunk, the block scans the chunk sums
__global__ __launch_bounds__(1024) void k_scan_start(int n, const int32_t *__restrict__ count,
                                                     int32_t *__restrict__ start) {
  __shared__ int s[1024];
  const int per = (n + 1023) / 1024;
  const int lo = min((int)threadIdx.x * per, n), hi = min(lo + per, n);
  int sum = 0;
  for (int i = lo; i < hi; i++) sum += count[i];
  s[threadIdx.x] = sum;
  __syncthreads();
  for (int off = 1; off < 1024; off <<= 1) {
    const int v = (threadIdx.x >= (unsigned)off) ? s[threadIdx.x - off] : 0;
    __syncthreads();
    s[threadIdx.x] += v;
    __syncthreads();
  }
  int run = s[threadIdx.x] - sum;
  for (int i = lo; i < hi; i++) {
    start[i] = run;
    run += count[i];
  }
  if (threadIdx.x == 1023) start[n] = s[1023];
}

// items of every bin in rank order; the counts (already scanned) are left
// zeroed for the next call's k_prep
__global__ __launch_bounds__(256) void k_tile_fill(int nitem, const int4 *__restrict__ base,
                                                   PmeGeom g, TileGeom tg,
                                                   int32_t *__restrict__ count,
                                                   const int32_t *__restrict__ start,
                                                   const int32_t *__restrict__ rank,
                                                   int32_t *__restrict__ items) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  for (int z = k; z < tg.ntile; z += gridDim.x * blockDim.x) count[z] = 0;
  if (k >= nitem) return;
  const int r = rank[k];
  if (r < 0) return;
  items[start[tile_home(base[k], g, tg)] + r] = k;
}

template <int PREC> struct SpreadAcc { using type = typename Prec<PREC>::real; };
template <> struct SpreadAcc<PGK_PREC_MIXED> { using type = unsigned long long; };

// the stencil points of one axis that fall in the tile [o, o + T): bit a set
// when point b + a (wrapped once into [0, K)) does
template <int P>
__device__ __forceinline__ unsigned stencil_bits(int b, int o, int T, int K) {
  unsigned m = 0;
#pragma unroll
  for (int a = 0; a < P; a++) {
    int i = b + a;
    if (i < 0) i += K;
    if ((unsigned)(i - o) < (unsigned)T) m |= 1u << a;
  }
  return m;
}

// One wave per tile of the charge grid (owner computes: no global atomics).
// The atoms that can reach the tile are those whose stencil origin lies in
// its home bin or in the nsrc - 1 bins below it along each axis.  The wave
// first keeps, of those visits, the ones whose stencil overlaps the tile at
// all (about a third: an atom of a lower bin reaches up only from its last
// P - 1 planes), then each lane takes one kept visit and adds exactly the
// stencil points inside the tile to the LDS tile.  Mixed form: the points
// are 2^-40 fixed-point integers, so the sums do not depend on the order of
// the additions; the tile is then written to the grid once.
template <int PREC, int P>
__global__ __launch_bounds__(64) void k_spread_tile(int nitem, const double *__restrict__ q,
                                                    PmeGeom g, TileGeom tg,
                                                    const int32_t *__restrict__ start,
                                                    const int32_t *__restrict__ items,
                                                    const typename Prec<PREC>::real *__restrict__ th,
                                                    const int4 *__restrict__ base,
                                                    typename Prec<PREC>::real *__restrict__ grid) {
  using R = typename Prec<PREC>::real;
  using S = typename SpreadAcc<PREC>::type;
  extern __shared__ __align__(16) unsigned char smem[];
  S *acc = reinterpret_cast<S *>(smem);
  __shared__ int s_lo[PGK_MAXSRC], s_n[PGK_MAXSRC + 1];
  __shared__ int s_vis[PGK_SPREAD_MAXVIS];
  __shared__ int s_nv;
  const int T0 = tg.T[0], T1 = tg.T[1], T2 = tg.T[2];
  const int npt = T0 * T1 * T2;
  const int tid = xcd_block();
  if (tid >= tg.ntile) return;
  const int lane = threadIdx.x;
  const int tx = tid % tg.nt[0], ty = (tid / tg.nt[0]) % tg.nt[1], tz = tid / (tg.nt[0] * tg.nt[1]);
  const int ox = tx * T0, oy = ty * T1, oz = tz * T2;
  for (int i = lane; i < npt; i += 64) acc[i] = S(0);
  // the source bins of this tile, flattened into one index space
  const int nsrc = tg.nsrc[0] * tg.nsrc[1] * tg.nsrc[2];
  for (int u = lane; u < nsrc; u += 64) {
    const int dx = u % tg.nsrc[0], dy = (u / tg.nsrc[0]) % tg.nsrc[1],
              dz = u / (tg.nsrc[0] * tg.nsrc[1]);
    const int sx = (tx - dx + tg.nt[0]) % tg.nt[0];
    const int sy = (ty - dy + tg.nt[1]) % tg.nt[1];
    const int sz = (tz - dz + tg.nt[2]) % tg.nt[2];
    const int bin = (sz * tg.nt[1] + sy) * tg.nt[0] + sx;
    s_lo[u] = start[bin];
    s_n[u + 1] = start[bin + 1] - start[bin];
  }
  if (lane == 0) s_nv = 0;
  __syncthreads();
  if (lane == 0) {
    s_n[0] = 0;
    for (int u = 0; u < nsrc; u++) s_n[u + 1] += s_n[u];
  }
  __syncthreads();
  const int total = s_n[nsrc];
  // keep the visits whose stencil meets the tile (the kept order does not
  // matter: the additions below commute exactly in the mixed form)
  for (int v = lane; v < total; v += 64) {
    int u = 0;
    while (u + 1 < nsrc && v >= s_n[u + 1]) u++;
    const int k = items[s_lo[u] + (v - s_n[u])];
    const int4 b = base[k];
    if (stencil_bits<P>(b.x, ox, T0, g.K[0]) && stencil_bits<P>(b.y, oy, T1, g.K[1]) &&
        stencil_bits<P>(b.z, oz, T2, g.K[2])) {
      const int w = atomicAdd(&s_nv, 1);
      if (w < PGK_SPREAD_MAXVIS) s_vis[w] = k;
    }
  }
  __syncthreads();
  const int nv = min(s_nv, PGK_SPREAD_MAXVIS);
  const size_t st = (size_t)nitem;
  // the rare tile with more kept visits than fit in the list takes the rest
  // with a second walk over all visits (same test, no list)
  const bool overflow = s_nv > PGK_SPREAD_MAXVIS;
  for (int w = lane; w < (overflow ? total + nv : nv); w += 64) {
    int k;
    if (w < nv) {
      k = s_vis[w];
    } else {
      const int v = w - nv;
      int u = 0;
      while (u + 1 < nsrc && v >= s_n[u + 1]) u++;
      k = items[s_lo[u] + (v - s_n[u])];
      bool listed = false;
      for (int x = 0; x < nv && !listed; x++) listed = s_vis[x] == k;
      if (listed) continue;
    }
    const int4 b = base[k];
    const unsigned mx = stencil_bits<P>(b.x, ox, T0, g.K[0]);
    const unsigned my = stencil_bits<P>(b.y, oy, T1, g.K[1]);
    const unsigned mz = stencil_bits<P>(b.z, oz, T2, g.K[2]);
    if (!(mx && my && mz)) continue;
    const R qi = (R)q[b.w];
    R w1[P], w2[P], w3[P];
#pragma unroll
    for (int a = 0; a < P; a++) {
      w1[a] = th[(size_t)a * st + k];
      w2[a] = th[(size_t)(P + a) * st + k];
      w3[a] = th[(size_t)(2 * P + a) * st + k] * qi;
    }
#pragma unroll
    for (int a3 = 0; a3 < P; a3++) {
      if (!((mz >> a3) & 1u)) continue;
      int iz = b.z + a3;
      if (iz < 0) iz += g.K[2];
      const int lz = iz - oz;
#pragma unroll
      for (int a2 = 0; a2 < P; a2++) {
        if (!((my >> a2) & 1u)) continue;
        int iy = b.y + a2;
        if (iy < 0) iy += g.K[1];
        const int ly = iy - oy;
        const R w23 = w2[a2] * w3[a3];
#pragma unroll
        for (int a1 = 0; a1 < P; a1++) {
          if (!((mx >> a1) & 1u)) continue;
          int ix = b.x + a1;
          if (ix < 0) ix += g.K[0];
          const R val = w1[a1] * w23;
          S *dst = &acc[(lz * T1 + ly) * T0 + (ix - ox)];
          if (PREC == PGK_PREC_MIXED)
            atomicAdd(dst, (S)fx_of(val * 256.0f));   // 2^-40 fixed point
          else
            atomicAdd(dst, (S)val);
        }
      }
    }
  }
  __syncthreads();
  for (int i = lane; i < npt; i += 64) {
    const int lx = i % T0, ly = (i / T0) % T1, lz = i / (T0 * T1);
    const size_t idx = ((size_t)(oz + lz) * g.K[1] + (oy + ly)) * g.K[0] + (ox + lx);
    if (PREC == PGK_PREC_MIXED)
      grid[idx] = (R)((double)(long long)acc[i] * PGK_GRID_FX_INV);
    else
      grid[idx] = (R)acc[i];
  }
}


// Global -> LDS staging: each thread issues PGK_STAGE loads before the first
// LDS store, so their memory latencies overlap instead of adding up.
#define PGK_STAGE 8
template <class V, class Ld, class St>
__device__ __forceinline__ void stage_in(int total, Ld ld, St st) {
  for (int t0 = threadIdx.x; t0 < total; t0 += PGK_STAGE * blockDim.x) {
    V v[PGK_STAGE];
#pragma unroll
    for (int u = 0; u < PGK_STAGE; u++) {
      const int t = t0 + u * blockDim.x;
      if (t < total) v[u] = ld(t);
    }
#pragma unroll
    for (int u = 0; u < PGK_STAGE; u++) {
      const int t = t0 + u * blockDim.x;
      if (t < total) st(t, v[u]);
    }
  }
}

// The half spectrum is stored [k3][h][k2] (k2 fastest): row r = k3 K2 + k2 of
// the real grid puts its mode h at cidx(r, h).  Lines along k2 are then
// contiguous, and lines along k3 and the x pass's row pairs read whole
// cache lines.
__device__ __forceinline__ size_t cidx(int r, int h, int H, int K2) {
  const int k3 = r / K2;
  return ((size_t)k3 * H + h) * K2 + (r - k3 * K2);
}

// x pass, forward: pairs of real rows -> half spectra.  One block handles NL
// complex lines (2*NL rows).
template <int PREC>
__global__ __launch_bounds__(256) void k_fft_x_fwd(const void *grid, void *cgrid, int nrow,
                                                    int NL, FftPlan p, const void *twv, int H,
                                                    int K2) {
  using R = typename Prec<PREC>::real;
  using C = typename C2T<R>::type;
  extern __shared__ __align__(16) unsigned char smem[];
  const int n = p.n;
  C *pa = reinterpret_cast<C *>(smem);
  C *pb = pa + NL * n;
  C *tw = pb + NL * n;
  for (int t = threadIdx.x; t < n; t += blockDim.x) tw[t] = reinterpret_cast<const C *>(twv)[t];
  const int row0 = blockIdx.x * 2 * NL;
  const float rn = 1.0f / (float)n;
  const R *gr = reinterpret_cast<const R *>(grid);
  stage_in<C>(
      NL * n,
      [&](int t) {
        const int l = fdiv(t, n, rn), k = t - l * n;
        const int ra = row0 + 2 * l, rb = ra + 1;
        C z;
        z.x = (ra < nrow) ? gr[(size_t)ra * n + k] : R(0);
        z.y = (rb < nrow) ? gr[(size_t)rb * n + k] : R(0);
        return z;
      },
      [&](int t, C z) { pa[t] = z; });
  __syncthreads();
  fft_lds<C>(pa, pb, NL, n, p, tw, -1);
  C *cg = reinterpret_cast<C *>(cgrid);
  const float rH = 1.0f / (float)H;
  for (int t = threadIdx.x; t < NL * H; t += blockDim.x) {
    int l = fdiv(t, H, rH), k = t - l * H;
    int ra = row0 + 2 * l, rb = ra + 1;
    C zk = pa[l * n + k];
    C zn = pa[l * n + ((n - k) % n)];
    // A = (Z_k + conj Z_{n-k})/2,  B = (Z_k - conj Z_{n-k})/(2i)
    C A, B;
    A.x = R(0.5) * (zk.x + zn.x);
    A.y = R(0.5) * (zk.y - zn.y);
    B.x = R(0.5) * (zk.y + zn.y);
    B.y = R(0.5) * (zn.x - zk.x);
    if (ra < nrow) cg[cidx(ra, k, H, K2)] = A;
    if (rb < nrow) cg[cidx(rb, k, H, K2)] = B;
  }
}

// x pass, backward: half spectra of two rows -> two real rows (C2R)
template <int PREC>
__global__ __launch_bounds__(256) void k_fft_x_bwd(const void *cgrid, void *grid, int nrow,
                                                    int NL, FftPlan p, const void *twv, int H,
                                                    int K2) {
  using R = typename Prec<PREC>::real;
  using C = typename C2T<R>::type;
  extern __shared__ __align__(16) unsigned char smem[];
  const int n = p.n;
  C *pa = reinterpret_cast<C *>(smem);
  C *pb = pa + NL * n;
  C *tw = pb + NL * n;
  for (int t = threadIdx.x; t < n; t += blockDim.x) tw[t] = reinterpret_cast<const C *>(twv)[t];
  const C *cg = reinterpret_cast<const C *>(cgrid);
  const int row0 = blockIdx.x * 2 * NL;
  const float rn = 1.0f / (float)n;
  stage_in<C>(
      NL * n,
      [&](int t) {
        const int l = fdiv(t, n, rn), k = t - l * n;
        const int ra = row0 + 2 * l, rb = ra + 1;
        C A, B;
        A.x = A.y = B.x = B.y = R(0);
        const int kk = (k < H) ? k : n - k;
        if (ra < nrow) A = cg[cidx(ra, kk, H, K2)];
        if (rb < nrow) B = cg[cidx(rb, kk, H, K2)];
        if (k >= H) {
          A.y = -A.y;
          B.y = -B.y;
        }
        // Z = A + i B
        C z;
        z.x = A.x - B.y;
        z.y = A.y + B.x;
        return z;
      },
      [&](int t, C z) { pa[t] = z; });
  __syncthreads();
  fft_lds<C>(pa, pb, NL, n, p, tw, +1);
  R *gr = reinterpret_cast<R *>(grid);
  for (int t = threadIdx.x; t < NL * n; t += blockDim.x) {
    int l = fdiv(t, n, rn), k = t - l * n;
    int ra = row0 + 2 * l, rb = ra + 1;
    C z = pa[t];
    if (ra < nrow) gr[(size_t)ra * n + k] = z.x;
    if (rb < nrow) gr[(size_t)rb * n + k] = z.y;
  }
}

// y pass (both directions): lines along k2 for a tile of TW consecutive h at one k3
template <int PREC>
__global__ __launch_bounds__(256) void k_fft_y(void *cgrid, int H, int K2, int TW, FftPlan p,
                                                const void *twv, int sgn) {
  using R = typename Prec<PREC>::real;
  using C = typename C2T<R>::type;
  extern __shared__ __align__(16) unsigned char smem[];
  const int n = p.n;   // == K2
  const int ls = n + 1;
  C *pa = reinterpret_cast<C *>(smem);
  C *pb = pa + TW * ls;
  C *tw = pb + TW * ls;
  for (int t = threadIdx.x; t < n; t += blockDim.x) tw[t] = reinterpret_cast<const C *>(twv)[t];
  C *cg = reinterpret_cast<C *>(cgrid);
  const int ntile = (H + TW - 1) / TW;
  const int k3 = blockIdx.x / ntile;
  const int h0 = (blockIdx.x - k3 * ntile) * TW;
  const int lt = __ffs(TW) - 1;   // TW is a power of two
  stage_in<C>(
      TW * n,
      [&](int t) {
        const int k2 = t >> lt, l = t & (TW - 1);
        C z;
        z.x = z.y = R(0);
        if (h0 + l < H) z = cg[((size_t)k3 * H + h0 + l) * K2 + k2];
        return z;
      },
      [&](int t, C z) { pa[(t & (TW - 1)) * ls + (t >> lt)] = z; });
  __syncthreads();
  fft_lds<C>(pa, pb, TW, ls, p, tw, sgn);
  for (int t = threadIdx.x; t < TW * n; t += blockDim.x) {
    int k2 = t >> lt, l = t & (TW - 1);
    if (h0 + l < H) cg[((size_t)k3 * H + h0 + l) * K2 + k2] = pa[l * ls + k2];
  }
}

// Small grids: the x and y passes of one z plane in one block, the plane
// held in LDS between them (forward: real rows -> x half spectra -> y lines;
// backward the reverse).  The arithmetic per line is the separate passes'
// (k_fft_x_fwd / k_fft_y / k_fft_x_bwd), so the results are the same; the
// reciprocal chain of a small system loses two launches per force call.
// LDS: row pairs [NL][K1] x 2, y lines [H][K2 + 1] x 2, twiddles K1 + K2.
template <int PREC>
__global__ __launch_bounds__(256) void k_fft_xy_fwd(const void *grid, void *cgrid, int K1,
                                                     int K2, int H, FftPlan p1, const void *tw1v,
                                                     FftPlan p2, const void *tw2v) {
  using R = typename Prec<PREC>::real;
  using C = typename C2T<R>::type;
  extern __shared__ __align__(16) unsigned char smem[];
  const int NL = (K2 + 1) / 2, ly = K2 + 1;
  C *pa = reinterpret_cast<C *>(smem);
  C *pb = pa + NL * K1;
  C *qa = pb + NL * K1;
  C *qb = qa + H * ly;
  C *tw1 = qb + H * ly;
  C *tw2 = tw1 + K1;
  for (int t = threadIdx.x; t < K1; t += blockDim.x) tw1[t] = reinterpret_cast<const C *>(tw1v)[t];
  for (int t = threadIdx.x; t < K2; t += blockDim.x) tw2[t] = reinterpret_cast<const C *>(tw2v)[t];
  const int k3 = blockIdx.x;
  const R *gr = reinterpret_cast<const R *>(grid) + (size_t)k3 * K2 * K1;
  const float r1 = 1.0f / (float)K1, rh = 1.0f / (float)H, r2 = 1.0f / (float)K2;
  for (int t = threadIdx.x; t < NL * K1; t += blockDim.x) {
    const int l = fdiv(t, K1, r1), k = t - l * K1;
    const int ra = 2 * l, rb = ra + 1;
    C z;
    z.x = gr[(size_t)ra * K1 + k];
    z.y = (rb < K2) ? gr[(size_t)rb * K1 + k] : R(0);
    pa[t] = z;
  }
  __syncthreads();
  fft_lds<C>(pa, pb, NL, K1, p1, tw1, -1);
  for (int t = threadIdx.x; t < NL * H; t += blockDim.x) {
    const int l = fdiv(t, H, rh), k = t - l * H;
    const int ra = 2 * l, rb = ra + 1;
    const C zk = pa[l * K1 + k];
    const C zn = pa[l * K1 + ((K1 - k) % K1)];
    // A = (Z_k + conj Z_{n-k})/2,  B = (Z_k - conj Z_{n-k})/(2i)
    C A, B;
    A.x = R(0.5) * (zk.x + zn.x);
    A.y = R(0.5) * (zk.y - zn.y);
    B.x = R(0.5) * (zk.y + zn.y);
    B.y = R(0.5) * (zn.x - zk.x);
    qa[k * ly + ra] = A;
    if (rb < K2) qa[k * ly + rb] = B;
  }
  __syncthreads();
  fft_lds<C>(qa, qb, H, ly, p2, tw2, -1);
  C *cg = reinterpret_cast<C *>(cgrid) + (size_t)k3 * H * K2;
  for (int t = threadIdx.x; t < H * K2; t += blockDim.x) {
    const int h = fdiv(t, K2, r2), k2 = t - h * K2;
    cg[(size_t)h * K2 + k2] = qa[h * ly + k2];
  }
}

template <int PREC>
__global__ __launch_bounds__(256) void k_fft_yx_bwd(const void *cgrid, void *grid, int K1,
                                                     int K2, int H, FftPlan p1, const void *tw1v,
                                                     FftPlan p2, const void *tw2v) {
  using R = typename Prec<PREC>::real;
  using C = typename C2T<R>::type;
  extern __shared__ __align__(16) unsigned char smem[];
  const int NL = (K2 + 1) / 2, ly = K2 + 1;
  C *pa = reinterpret_cast<C *>(smem);
  C *pb = pa + NL * K1;
  C *qa = pb + NL * K1;
  C *qb = qa + H * ly;
  C *tw1 = qb + H * ly;
  C *tw2 = tw1 + K1;
  for (int t = threadIdx.x; t < K1; t += blockDim.x) tw1[t] = reinterpret_cast<const C *>(tw1v)[t];
  for (int t = threadIdx.x; t < K2; t += blockDim.x) tw2[t] = reinterpret_cast<const C *>(tw2v)[t];
  const int k3 = blockIdx.x;
  const C *cg = reinterpret_cast<const C *>(cgrid) + (size_t)k3 * H * K2;
  const float r1 = 1.0f / (float)K1, r2 = 1.0f / (float)K2;
  for (int t = threadIdx.x; t < H * K2; t += blockDim.x) {
    const int h = fdiv(t, K2, r2), k2 = t - h * K2;
    qa[h * ly + k2] = cg[(size_t)h * K2 + k2];
  }
  __syncthreads();
  fft_lds<C>(qa, qb, H, ly, p2, tw2, +1);
  for (int t = threadIdx.x; t < NL * K1; t += blockDim.x) {
    const int l = fdiv(t, K1, r1), k = t - l * K1;
    const int ra = 2 * l, rb = ra + 1;
    const int kk = (k < H) ? k : K1 - k;
    C A = qa[kk * ly + ra], B;
    B.x = B.y = R(0);
    if (rb < K2) B = qa[kk * ly + rb];
    if (k >= H) {
      A.y = -A.y;
      B.y = -B.y;
    }
    // Z = A + i B
    C z;
    z.x = A.x - B.y;
    z.y = A.y + B.x;
    pa[t] = z;
  }
  __syncthreads();
  fft_lds<C>(pa, pb, NL, K1, p1, tw1, +1);
  R *gr = reinterpret_cast<R *>(grid) + (size_t)k3 * K2 * K1;
  for (int t = threadIdx.x; t < NL * K1; t += blockDim.x) {
    const int l = fdiv(t, K1, r1), k = t - l * K1;
    const int ra = 2 * l, rb = ra + 1;
    const C z = pa[t];
    gr[(size_t)ra * K1 + k] = z.x;
    if (rb < K2) gr[(size_t)rb * K1 + k] = z.y;
  }
}

// y pass with the register FFT: one wavefront per block, LPW lines (h fastest
// across lines, then k3), so consecutive lanes read consecutive h
template <int PREC, class F, int SGN>
__global__ __launch_bounds__(64) void k_fft_y_reg(void *cgrid, int H, int K3,
                                                  const void *twv) {
  using R = typename Prec<PREC>::real;
  using C = typename C2T<R>::type;
  __shared__ C tw[F::N];
  __shared__ C ex[F::BUF];
  const int lane = threadIdx.x, t = lane % F::N2, l = lane / F::N2;
  for (int k = lane; k < F::N; k += 64) tw[k] = reinterpret_cast<const C *>(twv)[k];
  // line (k3, h) = k3 H + h is contiguous along k2
  const int line = blockIdx.x * F::LPW + l;
  const bool valid = line < H * K3;
  C *cg = reinterpret_cast<C *>(cgrid);
  const size_t base = (size_t)(valid ? line : 0) * F::N;
  C x[F::N1];
#pragma unroll
  for (int j = 0; j < F::N1; j++) {
    x[j].x = x[j].y = R(0);
    if (valid) x[j] = cg[base + t + F::N2 * j];
  }
  __syncthreads();
  fft_reg<F, SGN>(x, ex, tw, t, l);
  if (valid) {
#pragma unroll
    for (int q = 0; q < F::N1 / F::N2; q++)
#pragma unroll
      for (int k2 = 0; k2 < F::N2; k2++)
        cg[base + t + F::N2 * q + F::N1 * k2] = x[q * F::N2 + k2];
  }
}

// x pass, forward, register FFT: complex line p packs real rows 2p and 2p+1
template <int PREC, class F>
__global__ __launch_bounds__(64) void k_fft_x_fwd_reg(const void *grid, void *cgrid, int nrow,
                                                      const void *twv, int H, int K2) {
  using R = typename Prec<PREC>::real;
  using C = typename C2T<R>::type;
  __shared__ C tw[F::N];
  __shared__ C ex[F::BUF];
  const int lane = threadIdx.x, t = lane % F::N2, l = lane / F::N2;
  for (int k = lane; k < F::N; k += 64) tw[k] = reinterpret_cast<const C *>(twv)[k];
  const int ra = 2 * (blockIdx.x * F::LPW + l), rb = ra + 1;
  const R *gr = reinterpret_cast<const R *>(grid);
  C x[F::N1];
#pragma unroll
  for (int j = 0; j < F::N1; j++) {
    const int k = t + F::N2 * j;
    x[j].x = (ra < nrow) ? gr[(size_t)ra * F::N + k] : R(0);
    x[j].y = (rb < nrow) ? gr[(size_t)rb * F::N + k] : R(0);
  }
  __syncthreads();
  fft_reg<F, -1>(x, ex, tw, t, l);
  // natural order, then split the two real rows' spectra
#pragma unroll
  for (int q = 0; q < F::N1 / F::N2; q++)
#pragma unroll
    for (int k2 = 0; k2 < F::N2; k2++)
      ex[l * F::NS + t + F::N2 * q + F::N1 * k2] = x[q * F::N2 + k2];
  __syncthreads();
  C *cg = reinterpret_cast<C *>(cgrid);
  for (int k = t; k < H; k += F::N2) {
    const C zk = ex[l * F::NS + k];
    const C zn = ex[l * F::NS + ((F::N - k) & (F::N - 1))];
    // A = (Z_k + conj Z_{n-k})/2,  B = (Z_k - conj Z_{n-k})/(2i)
    C A, B;
    A.x = R(0.5) * (zk.x + zn.x);
    A.y = R(0.5) * (zk.y - zn.y);
    B.x = R(0.5) * (zk.y + zn.y);
    B.y = R(0.5) * (zn.x - zk.x);
    if (ra < nrow) cg[cidx(ra, k, H, K2)] = A;
    if (rb < nrow) cg[cidx(rb, k, H, K2)] = B;
  }
}

// x pass, backward (C2R), register FFT
template <int PREC, class F>
__global__ __launch_bounds__(64) void k_fft_x_bwd_reg(const void *cgrid, void *grid, int nrow,
                                                      const void *twv, int H, int K2) {
  using R = typename Prec<PREC>::real;
  using C = typename C2T<R>::type;
  __shared__ C tw[F::N];
  __shared__ C ex[F::BUF];
  const int lane = threadIdx.x, t = lane % F::N2, l = lane / F::N2;
  for (int k = lane; k < F::N; k += 64) tw[k] = reinterpret_cast<const C *>(twv)[k];
  const int ra = 2 * (blockIdx.x * F::LPW + l), rb = ra + 1;
  const C *cg = reinterpret_cast<const C *>(cgrid);
  C x[F::N1];
#pragma unroll
  for (int j = 0; j < F::N1; j++) {
    const int k = t + F::N2 * j;
    const int kk = (k < H) ? k : F::N - k;
    C A, B;
    A.x = A.y = B.x = B.y = R(0);
    if (ra < nrow) A = cg[cidx(ra, kk, H, K2)];
    if (rb < nrow) B = cg[cidx(rb, kk, H, K2)];
    if (k >= H) {
      A.y = -A.y;
      B.y = -B.y;
    }
    // Z = A + i B
    x[j].x = A.x - B.y;
    x[j].y = A.y + B.x;
  }
  __syncthreads();
  fft_reg<F, +1>(x, ex, tw, t, l);
  R *gr = reinterpret_cast<R *>(grid);
#pragma unroll
  for (int q = 0; q < F::N1 / F::N2; q++)
#pragma unroll
    for (int k2 = 0; k2 < F::N2; k2++) {
      const int k = t + F::N2 * q + F::N1 * k2;
      if (ra < nrow) gr[(size_t)ra * F::N + k] = x[q * F::N2 + k2].x;
      if (rb < nrow) gr[(size_t)rb * F::N + k] = x[q * F::N2 + k2].y;
    }
}

struct ConvArgs {
  const double *bsp;    // K1 | K2 | K3 moduli
  double recip[3];
  double fac, rk2, pivol, rpivol, netq;
  int K[3], H;
  int ifnetq;
};

// Influence function of mode (h, k2, k3) (scalar_sum, fft_pme.F90:544-690):
// returns eterm (0 outside |m| <= rkcut) and adds the mode's energy and
// virial terms, weighted for the half spectrum, to ev.  mh2 = m2/L2 and s2
// (+1 on the k2 Nyquist plane, else -1) belong to k2; b1 b2 b3 are the
// B-spline moduli of h, k2, k3.
__device__ __forceinline__ double conv_mode(const ConvArgs &ca, int h, int k3, double mh2,
                                            double s2, double b1, double b2, double b3,
                                            double zx, double zy, double (&ev)[8]) {
  const int K1 = ca.K[0], K3 = ca.K[2];
  const int nf1 = (K1 + 1) / 2, nf3 = (K3 + 1) / 2;
  const int m1 = (h < nf1) ? h : h - K1;
  const int m3 = (k3 < nf3) ? k3 : k3 - K3;
  const double mh1 = ca.recip[0] * m1, mh3 = ca.recip[2] * m3;
  // the rkcut test decides modes that lie on the sphere (e.g. integer
  // |k|^2 = (rkcut L)^2): |m|^2 in the reference's rounding, no contraction
  const double msq = msq_nc(mh1, mh2, mh3);
  double eterm = 0.0;
  if (msq <= ca.rk2) {
    const double denom = ca.pivol * b1 * b2 * b3 * msq;
    const double ex = exp(-ca.fac * msq);
    eterm = ex / denom;
    const double vterm = 2.0 * (ca.fac * msq + 1.0) / msq;
    const double struc2 = zx * zx + zy * zy;
    const bool single = (h == 0) || ((K1 & 1) == 0 && h == K1 / 2);
    const double w = single ? 1.0 : 2.0;
    const double s3 = ((K3 & 1) == 0 && k3 == K3 / 2) ? 1.0 : -1.0;
    const double w12 = single ? 1.0 : 1.0 - s2;        // s1 = -1
    const double w13 = single ? 1.0 : 1.0 - s3;
    const double w23 = single ? 1.0 : 1.0 + s2 * s3;
    double wgt = eterm * struc2;
    if (ca.ifnetq) {
      const double e2 = ex * ca.netq / (ca.pivol * msq);
      ev[1] += w * (e2 * (-1.0 + 2.0 * ca.fac * msq));
      wgt = wgt - e2;
    }
    ev[0] += w * eterm * struc2;
    ev[2] += w * wgt * (vterm * mh1 * mh1 - 1.0);
    ev[3] += w * wgt * (vterm * mh2 * mh2 - 1.0);
    ev[4] += w * wgt * (vterm * mh3 * mh3 - 1.0);
    ev[5] += w12 * wgt * (vterm * mh1 * mh2);
    ev[6] += w13 * wgt * (vterm * mh1 * mh3);
    ev[7] += w23 * wgt * (vterm * mh2 * mh3);
  }
  return eterm;
}

// The same mode with exp(-fac |m|^2) taken as the product of per-axis factors
// exp(-fac mh_d^2) and the moduli as reciprocals, so that a mode costs one
// division.  e12 = exp(-fac (mh1^2 + mh2^2)) as a product, r12 = 1/(pi V b1 b2),
// e3 = exp(-fac mh3^2), r3 = 1/b3; the cutoff test is the reference's.
__device__ __forceinline__ double conv_mode_f(const ConvArgs &ca, int h, int k3, double mh1,
                                              double mh2, double s2, double e12, double r12,
                                              double e3, double r3, double zx, double zy,
                                              double (&ev)[8]) {
  const int K1 = ca.K[0], K3 = ca.K[2];
  const int nf3 = (K3 + 1) / 2;
  const int m3 = (k3 < nf3) ? k3 : k3 - K3;
  const double mh3 = ca.recip[2] * m3;
  const double msq = msq_nc(mh1, mh2, mh3);
  double eterm = 0.0;
  if (msq <= ca.rk2) {
    const double rm = 1.0 / msq;
    const double ex = e12 * e3;
    eterm = ex * (r12 * r3) * rm;
    const double vterm = 2.0 * (ca.fac * msq + 1.0) * rm;
    const double struc2 = zx * zx + zy * zy;
    const bool single = (h == 0) || ((K1 & 1) == 0 && h == K1 / 2);
    const double w = single ? 1.0 : 2.0;
    const double s3 = ((K3 & 1) == 0 && k3 == K3 / 2) ? 1.0 : -1.0;
    const double w12 = single ? 1.0 : 1.0 - s2;        // s1 = -1
    const double w13 = single ? 1.0 : 1.0 - s3;
    const double w23 = single ? 1.0 : 1.0 + s2 * s3;
    double wgt = eterm * struc2;
    if (ca.ifnetq) {
      const double e2 = ex * ca.netq * ca.rpivol * rm;
      ev[1] += w * (e2 * (-1.0 + 2.0 * ca.fac * msq));
      wgt = wgt - e2;
    }
    ev[0] += w * eterm * struc2;
    ev[2] += w * wgt * (vterm * mh1 * mh1 - 1.0);
    ev[3] += w * wgt * (vterm * mh2 * mh2 - 1.0);
    ev[4] += w * wgt * (vterm * mh3 * mh3 - 1.0);
    ev[5] += w12 * wgt * (vterm * mh1 * mh2);
    ev[6] += w13 * wgt * (vterm * mh1 * mh3);
    ev[7] += w23 * wgt * (vterm * mh2 * mh3);
  }
  return eterm;
}

// z pass with the register FFT and the convolution between the forward and
// the backward transform; LPW lines per wavefront, h fastest across lines
template <int PREC, class F>
__global__ __launch_bounds__(64) void k_fft_z_conv_reg(void *cgrid, const void *twv, ConvArgs ca,
                                                       double *red) {
  using R = typename Prec<PREC>::real;
  using C = typename C2T<R>::type;
  __shared__ C tw[F::N];
  __shared__ C ex[F::BUF];
  __shared__ double se3[F::N], sr3[F::N];
  const int lane = threadIdx.x, t = lane % F::N2, l = lane / F::N2;
  const int H = ca.H, K2 = ca.K[1], K1 = ca.K[0];
  for (int k = lane; k < F::N; k += 64) {
    tw[k] = reinterpret_cast<const C *>(twv)[k];
    const int m3 = (k < (F::N + 1) / 2) ? k : k - F::N;
    const double mh3 = ca.recip[2] * m3;
    se3[k] = exp(-ca.fac * (mh3 * mh3));
    sr3[k] = 1.0 / ca.bsp[K1 + K2 + k];
  }
  // line (h, k2) = h K2 + k2: neighbouring lanes hold neighbouring k2
  const int line = blockIdx.x * F::LPW + l;
  const bool valid = line < H * K2;
  const int h = valid ? line / K2 : 0, k2 = valid ? line - h * K2 : 0;
  C *cg = reinterpret_cast<C *>(cgrid);
  const size_t zstride = (size_t)K2 * H;
  const size_t base = (size_t)(valid ? line : 0);
  C x[F::N1];
#pragma unroll
  for (int j = 0; j < F::N1; j++) {
    x[j].x = x[j].y = R(0);
    if (valid) x[j] = cg[base + (size_t)(t + F::N2 * j) * zstride];
  }
  __syncthreads();
  fft_reg<F, -1>(x, ex, tw, t, l);
  double ev[8] = {0, 0, 0, 0, 0, 0, 0, 0};   // E, vsca, xx yy zz xy xz yz
  const int nf1 = (K1 + 1) / 2, nf2 = (K2 + 1) / 2;
  const int m1 = (h < nf1) ? h : h - K1;
  const int m2 = (k2 < nf2) ? k2 : k2 - K2;
  const double mh1 = ca.recip[0] * m1, mh2 = ca.recip[1] * m2;
  const double s2 = ((K2 & 1) == 0 && k2 == K2 / 2) ? 1.0 : -1.0;
  const double b1 = valid ? ca.bsp[h] : 1.0, b2 = valid ? ca.bsp[K1 + k2] : 1.0;
  const double e12 = exp(-ca.fac * (mh1 * mh1)) * exp(-ca.fac * (mh2 * mh2));
  const double r12 = 1.0 / (ca.pivol * b1 * b2);
#pragma unroll
  for (int q = 0; q < F::N1 / F::N2; q++)
#pragma unroll
    for (int kk = 0; kk < F::N2; kk++) {
      const int k3 = t + F::N2 * q + F::N1 * kk;
      C &z = x[q * F::N2 + kk];
      if (valid && !(h == 0 && k2 == 0 && k3 == 0)) {   // m = 0 left untouched
        const double eterm = conv_mode_f(ca, h, k3, mh1, mh2, s2, e12, r12, se3[k3], sr3[k3],
                                         (double)z.x, (double)z.y, ev);
        z.x = (R)(eterm * (double)z.x);
        z.y = (R)(eterm * (double)z.y);
      }
      // natural order for the inverse transform's input layout
      ex[k3 * F::LPW + l] = z;
    }
  __syncthreads();
#pragma unroll
  for (int j = 0; j < F::N1; j++) x[j] = ex[(t + F::N2 * j) * F::LPW + l];
  __syncthreads();
  fft_reg<F, +1>(x, ex, tw, t, l);
  if (valid) {
#pragma unroll
    for (int q = 0; q < F::N1 / F::N2; q++)
#pragma unroll
      for (int kk = 0; kk < F::N2; kk++)
        cg[base + (size_t)(t + F::N2 * q + F::N1 * kk) * zstride] = x[q * F::N2 + kk];
  }
  block_record<8, 1>(ev, red + (size_t)blockIdx.x * PGK_REC);
}

// z pass with the influence-function convolution between the forward and the
// backward transform (scalar_sum, fft_pme.F90:544-690)
template <int PREC>
__global__ __launch_bounds__(256) void k_fft_z_conv(void *cgrid, int TW, FftPlan p,
                                                     const void *twv, ConvArgs ca,
                                                     double *red) {
  using R = typename Prec<PREC>::real;
  using C = typename C2T<R>::type;
  extern __shared__ __align__(16) unsigned char smem[];
  const int n = p.n;   // == K3
  const int ls = n + 1;
  const int H = ca.H, K2 = ca.K[1];
  C *pa = reinterpret_cast<C *>(smem);
  C *pb = pa + TW * ls;
  C *tw = pb + TW * ls;
  for (int t = threadIdx.x; t < n; t += blockDim.x) tw[t] = reinterpret_cast<const C *>(twv)[t];
  C *cg = reinterpret_cast<C *>(cgrid);
  const int ntile = (H + TW - 1) / TW;
  const int k2 = blockIdx.x / ntile;
  const int h0 = (blockIdx.x - k2 * ntile) * TW;
  const size_t zstride = (size_t)K2 * H;
  const int lt = __ffs(TW) - 1;   // TW is a power of two
  // B-spline moduli of the tile's h, and of every k3 (k2 is fixed)
  double *sb1 = reinterpret_cast<double *>(tw + n);
  double *sb3 = sb1 + TW;
  for (int t = threadIdx.x; t < TW + n; t += blockDim.x) {
    if (t < TW) sb1[t] = (h0 + t < H) ? ca.bsp[h0 + t] : 1.0;
    else sb3[t - TW] = ca.bsp[ca.K[0] + K2 + (t - TW)];
  }
  stage_in<C>(
      TW * n,
      [&](int t) {
        const int k3 = t >> lt, l = t & (TW - 1);
        C z;
        z.x = z.y = R(0);
        if (h0 + l < H) z = cg[(size_t)k3 * zstride + (size_t)(h0 + l) * K2 + k2];
        return z;
      },
      [&](int t, C z) { pa[(t & (TW - 1)) * ls + (t >> lt)] = z; });
  __syncthreads();
  fft_lds<C>(pa, pb, TW, ls, p, tw, -1);
  // convolution, energy and virial
  double ev[8] = {0, 0, 0, 0, 0, 0, 0, 0};   // E, vsca, xx yy zz xy xz yz
  const int K1 = ca.K[0];
  const int nf2 = (K2 + 1) / 2;
  const int m2 = (k2 < nf2) ? k2 : k2 - K2;
  const double mh2 = ca.recip[1] * m2;
  const double s2 = ((K2 & 1) == 0 && k2 == K2 / 2) ? 1.0 : -1.0;
  const double bsp2 = ca.bsp[K1 + k2];
  const float rn = 1.0f / (float)n;
  for (int t = threadIdx.x; t < TW * n; t += blockDim.x) {
    int l = fdiv(t, n, rn), k3 = t - l * n;
    int h = h0 + l;
    if (h >= H) continue;
    if (h == 0 && k2 == 0 && k3 == 0) continue;   // m = 0 left untouched
    C z = pa[l * ls + k3];
    const double eterm = conv_mode(ca, h, k3, mh2, s2, sb1[l], bsp2, sb3[k3], (double)z.x,
                                   (double)z.y, ev);
    z.x = (R)(eterm * (double)z.x);
    z.y = (R)(eterm * (double)z.y);
    pa[l * ls + k3] = z;
  }
  __syncthreads();
  fft_lds<C>(pa, pb, TW, ls, p, tw, +1);
  for (int t = threadIdx.x; t < TW * n; t += blockDim.x) {
    int k3 = t >> lt, l = t & (TW - 1);
    if (h0 + l < H) cg[(size_t)k3 * zstride + (size_t)(h0 + l) * K2 + k2] = pa[l * ls + k3];
  }
  block_record<8, 4>(ev, red + (size_t)blockIdx.x * PGK_REC);
}

// Per-atom reciprocal virial, component comp (0..5: xx yy zz xy xz yz) of
// the tensor kernel G_ab(m) = eterm(m) (delta_ab - vterm m_a m_b), the
// virial of scalar_sum (fft_pme.F90:544-690) in the sign of the library's
// Coulomb virial tensor and of ewaldkp_hf's virielct_atm: the z pass of the transformed
// charge grid forward, the product with G_ab, and back, as k_fft_z_conv
// does with eterm.  Off-diagonal terms on a Nyquist plane that the
// reference's half-spectrum weights drop (w12, w13, w23 = 0) are zeroed, so
// the per-atom tensors add up to the reference's virial.
template <int PREC>
__global__ __launch_bounds__(256) void k_fft_z_vir(void *cgrid, int TW, FftPlan p,
                                                    const void *twv, ConvArgs ca, int comp) {
  using R = typename Prec<PREC>::real;
  using C = typename C2T<R>::type;
  extern __shared__ __align__(16) unsigned char smem[];
  const int n = p.n;   // == K3
  const int ls = n + 1;
  const int H = ca.H, K2 = ca.K[1];
  C *pa = reinterpret_cast<C *>(smem);
  C *pb = pa + TW * ls;
  C *tw = pb + TW * ls;
  for (int t = threadIdx.x; t < n; t += blockDim.x) tw[t] = reinterpret_cast<const C *>(twv)[t];
  C *cg = reinterpret_cast<C *>(cgrid);
  const int ntile = (H + TW - 1) / TW;
  const int k2 = blockIdx.x / ntile;
  const int h0 = (blockIdx.x - k2 * ntile) * TW;
  const size_t zstride = (size_t)K2 * H;
  const int lt = __ffs(TW) - 1;   // TW is a power of two
  double *sb1 = reinterpret_cast<double *>(tw + n);
  double *sb3 = sb1 + TW;
  for (int t = threadIdx.x; t < TW + n; t += blockDim.x) {
    if (t < TW) sb1[t] = (h0 + t < H) ? ca.bsp[h0 + t] : 1.0;
    else sb3[t - TW] = ca.bsp[ca.K[0] + K2 + (t - TW)];
  }
  stage_in<C>(
      TW * n,
      [&](int t) {
        const int k3 = t >> lt, l = t & (TW - 1);
        C z;
        z.x = z.y = R(0);
        if (h0 + l < H) z = cg[(size_t)k3 * zstride + (size_t)(h0 + l) * K2 + k2];
        return z;
      },
      [&](int t, C z) { pa[(t & (TW - 1)) * ls + (t >> lt)] = z; });
  __syncthreads();
  fft_lds<C>(pa, pb, TW, ls, p, tw, -1);
  const int K1 = ca.K[0], K3 = ca.K[2];
  const int nf1 = (K1 + 1) / 2, nf2 = (K2 + 1) / 2, nf3 = (K3 + 1) / 2;
  const int m2 = (k2 < nf2) ? k2 : k2 - K2;
  const double mh2 = ca.recip[1] * m2;
  const double s2 = ((K2 & 1) == 0 && k2 == K2 / 2) ? 1.0 : -1.0;
  const double bsp2 = ca.bsp[K1 + k2];
  const int a = comp < 3 ? comp : (comp == 5 ? 1 : 0);
  const int b = comp < 3 ? comp : (comp == 3 ? 1 : 2);
  const float rn = 1.0f / (float)n;
  for (int t = threadIdx.x; t < TW * n; t += blockDim.x) {
    int l = fdiv(t, n, rn), k3 = t - l * n;
    int h = h0 + l;
    if (h >= H) continue;
    double g = 0.0;
    if (h != 0 || k2 != 0 || k3 != 0) {
      double ev[8] = {0, 0, 0, 0, 0, 0, 0, 0};
      const double eterm = conv_mode(ca, h, k3, mh2, s2, sb1[l], bsp2, sb3[k3], 0.0, 0.0, ev);
      if (eterm != 0.0) {
        const int m1 = (h < nf1) ? h : h - K1;
        const int m3 = (k3 < nf3) ? k3 : k3 - K3;
        const double mh1 = ca.recip[0] * m1, mh3 = ca.recip[2] * m3;
        const double msq = msq_nc(mh1, mh2, mh3);
        const double vterm = 2.0 * (ca.fac * msq + 1.0) / msq;
        const double ma = a == 0 ? mh1 : (a == 1 ? mh2 : mh3);
        const double mb = b == 0 ? mh1 : (b == 1 ? mh2 : mh3);
        g = eterm * ((a == b ? 1.0 : 0.0) - vterm * ma * mb);
        const bool single = (h == 0) || ((K1 & 1) == 0 && h == K1 / 2);
        const double s3 = ((K3 & 1) == 0 && k3 == K3 / 2) ? 1.0 : -1.0;
        if (!single && ((comp == 3 && s2 > 0.0) || (comp == 4 && s3 > 0.0) ||
                        (comp == 5 && s2 * s3 < 0.0)))
          g = 0.0;
      }
    }
    C z = pa[l * ls + k3];
    z.x = (R)(g * (double)z.x);
    z.y = (R)(g * (double)z.y);
    pa[l * ls + k3] = z;
  }
  __syncthreads();
  fft_lds<C>(pa, pb, TW, ls, p, tw, +1);
  for (int t = threadIdx.x; t < TW * n; t += blockDim.x) {
    int k3 = t >> lt, l = t & (TW - 1);
    if (h0 + l < H) cg[(size_t)k3 * zstride + (size_t)(h0 + l) * K2 + k2] = pa[l * ls + k3];
  }
}

// the six per-atom components into the Fortran (3,3,natom) layout
__global__ __launch_bounds__(256) void k_vir_assemble(int n, const double *__restrict__ v6,
                                                      double *__restrict__ w) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const double xx = v6[i], yy = v6[n + i], zz = v6[2 * (size_t)n + i];
  const double xy = v6[3 * (size_t)n + i], xz = v6[4 * (size_t)n + i],
               yz = v6[5 * (size_t)n + i];
  double *o = w + 9 * (size_t)i;
  o[0] = xx; o[1] = xy; o[2] = xz;
  o[3] = xy; o[4] = yy; o[5] = yz;
  o[6] = xz; o[7] = yz; o[8] = zz;
}

FftPlan make_plan(int n) {
  FftPlan p;
  p.n = n;
  p.nrad = 0;
  int m = n;
  while (m % 4 == 0) { p.rad[p.nrad++] = 4; m /= 4; }
  while (m % 2 == 0) { p.rad[p.nrad++] = 2; m /= 2; }
  while (m % 3 == 0) { p.rad[p.nrad++] = 3; m /= 3; }
  while (m % 5 == 0) { p.rad[p.nrad++] = 5; m /= 5; }
  if (m != 1) p.nrad = -1;
  return p;
}

// calls fn(RegFft<N1, N2>{}) for line lengths with a register FFT
template <class Fn>
bool with_reg_fft(int n, Fn fn) {
  switch (n) {
    case 32: fn(RegFft<8, 4>{}); return true;
    case 64: fn(RegFft<8, 8>{}); return true;
    case 128: fn(RegFft<16, 8>{}); return true;
    case 256: fn(RegFft<16, 16>{}); return true;
    default: return false;
  }
}

// lines per block: two line buffers of stride n + 1 and the twiddles fit in 64 KiB
int lines_for(int n, size_t csize, int maxl) {
  int nl = maxl;
  while (nl > 1 && (size_t)(2 * nl * (n + 1) + n) * csize > 65536) nl >>= 1;
  return nl;
}
size_t fft_lds_bytes(int nl, int n, size_t csize) { return (size_t)(2 * nl * (n + 1) + n) * csize; }

// LDS of the fused plane kernels (k_fft_xy_fwd / k_fft_yx_bwd), 0 when they
// do not apply: a register FFT on x or y, or a plane above 64 KiB
size_t fft_plane_bytes(int K1, int K2, size_t csize) {
  if (with_reg_fft(K1, [](auto) {}) || with_reg_fft(K2, [](auto) {})) return 0;
  const int NL = (K2 + 1) / 2, H = K1 / 2 + 1;
  const size_t b = (size_t)(2 * NL * K1 + 2 * H * (K2 + 1) + K1 + K2) * csize;
  return b <= 65536 ? b : 0;
}

// Device layout of the per-slot PME data: slots are the list's spatial order
// when a list exists (ncl*8 with padding), else atom order.
struct PmeLayout {
  PmeGeom g;
  TileGeom tg;
  int nitem;
  const int32_t *order;
  void *th;
  int4 *base;
  int32_t *count, *start, *rank, *items;
};

static int pme_layout(pgk_ctx *c, size_t rsize, PmeLayout *L) {
  for (int d = 0; d < 3; d++) {
    L->g.K[d] = c->nfft[d];
    L->g.recip[d] = 1.0 / c->box[d];
  }
  L->g.H = c->nfft[0] / 2 + 1;
  L->g.order = c->order;
  TileGeom &tg = L->tg;
  tg.ntile = 1;
  // spreading tiles of up to 8^3 points (one wave each); small grids take
  // 4^3 tiles, whose eight times as many waves fill the device
  const int tmax = ((int64_t)c->nfft[0] * c->nfft[1] * c->nfft[2] <= 64 * 64 * 64) ? 4 : 8;
  for (int d = 0; d < 3; d++) {
    int T = tmax;
    while (c->nfft[d] % T) T--;
    tg.T[d] = T;
    tg.nt[d] = c->nfft[d] / T;
    const int nb = (T + c->order - 2) / T;
    tg.nsrc[d] = std::min(nb + 1, tg.nt[d]);
    tg.ntile *= tg.nt[d];
  }
  if (tg.nsrc[0] * tg.nsrc[1] * tg.nsrc[2] > PGK_MAXSRC) {
    pgk_set_error("SPME grid tiles too small for this pme_order");
    return PGK_EINVAL;
  }
  const bool sorted = c->have_list && c->d_cl_atom;
  L->nitem = sorted ? c->ncl * PGK_CL : c->natom;
  L->order = sorted ? c->d_cl_atom : nullptr;
  const size_t nitem = (size_t)L->nitem;
  size_t need = nitem * 3 * c->order * rsize + nitem * sizeof(int4);
  if (need > c->theta_bytes) {
    if (c->d_theta) (void)hipFree(c->d_theta);
    c->d_theta = nullptr;
    c->theta_bytes = 0;
    HIP_CHECK(hipMalloc(&c->d_theta, need));
    c->theta_bytes = need;
  }
  L->th = c->d_theta;
  L->base = reinterpret_cast<int4 *>(reinterpret_cast<char *>(c->d_theta) +
                                     nitem * 3 * c->order * rsize);
  // bins: count[ntile] | start[ntile + 1] | rank[nitem] | items[nitem]
  size_t tneed = sizeof(int32_t) * (2 * (size_t)tg.ntile + 1 + 2 * nitem);
  if (tneed > c->tile_bytes || c->tile_ntile != tg.ntile) {
    if (c->d_tile) (void)hipFree(c->d_tile);
    c->d_tile = nullptr;
    c->tile_bytes = 0;
    HIP_CHECK(hipMalloc(&c->d_tile, tneed));
    c->tile_bytes = tneed;
    c->tile_ntile = tg.ntile;
    // counts start at zero; afterwards k_tile_fill leaves them zeroed
    HIP_CHECK(hipMemsetAsync(c->d_tile, 0, sizeof(int32_t) * tg.ntile, c->stream));
  }
  L->count = reinterpret_cast<int32_t *>(c->d_tile);
  L->start = L->count + tg.ntile;
  L->rank = L->start + tg.ntile + 1;
  L->items = L->rank + nitem;
  return PGK_OK;
}

template <int PREC, bool PACK>
int launch_prep(pgk_ctx *c) {
  using R = typename Prec<PREC>::real;
  using R4 = typename Prec<PREC>::real4;
  PmeLayout L;
  int rc = pme_layout(c, sizeof(R), &L);
  if (rc) return rc;
  const int gb = pgk_grid_blocks(L.nitem, 256);
  R *th = reinterpret_cast<R *>(L.th);
  R4 *xq = reinterpret_cast<R4 *>(c->d_xq);
#define PGK_PREP(P)                                                                          \
  k_prep<PREC, P, PACK><<<gb, 256, 0, c->stream>>>(L.nitem, L.order, c->d_pos, c->d_q,      \
                                                   c->d_type, L.g, L.tg, th, L.base, L.count, \
                                                   L.rank, xq, c->d_ct)
  switch (c->order) {
    case 3: PGK_PREP(3); break;
    case 4: PGK_PREP(4); break;
    case 5: PGK_PREP(5); break;
    case 6: PGK_PREP(6); break;
    case 7: PGK_PREP(7); break;
    case 8: PGK_PREP(8); break;
    case 9: PGK_PREP(9); break;
    default: PGK_PREP(10); break;
  }
#undef PGK_PREP
  HIP_CHECK(hipGetLastError());
  return PGK_OK;
}

template <int PREC>
int launch_pme(pgk_ctx *c, bool split, int *off, int *nrec) {
  using R = typename Prec<PREC>::real;
  using C = typename C2T<R>::type;
  using A = typename Prec<PREC>::acc;
  hipStream_t st = c->stream;
  const int K1 = c->nfft[0], K2 = c->nfft[1], K3 = c->nfft[2];
  const int H = K1 / 2 + 1;
  PmeLayout L;
  int rc = pme_layout(c, sizeof(R), &L);
  if (rc) return rc;
  const PmeGeom g = L.g;
  const TileGeom tg = L.tg;
  const int nitem = L.nitem;
  R *th = reinterpret_cast<R *>(L.th);
  int4 *base = L.base;
  pgk_prof_begin(c, PGK_PROF_SPREAD);
  {
    k_scan_start<<<1, 1024, 0, st>>>(tg.ntile, L.count, L.start);
    k_tile_fill<<<pgk_grid_blocks(std::max(nitem, tg.ntile), 256), 256, 0, st>>>(
        nitem, base, g, tg, L.count, L.start, L.rank, L.items);
    const size_t lds = (size_t)tg.T[0] * tg.T[1] * tg.T[2] * sizeof(typename SpreadAcc<PREC>::type);
    R *gr = reinterpret_cast<R *>(c->d_grid);
    const int32_t *start = L.start, *items = L.items;
#define PGK_SPREAD(P)                                                                         \
  k_spread_tile<PREC, P><<<xcd_grid(tg.ntile), 64, lds, st>>>(nitem, c->d_q, g, tg, start,    \
                                                                items, th, base, gr)
    switch (c->order) {
      case 3: PGK_SPREAD(3); break;
      case 4: PGK_SPREAD(4); break;
      case 5: PGK_SPREAD(5); break;
      case 6: PGK_SPREAD(6); break;
      case 7: PGK_SPREAD(7); break;
      case 8: PGK_SPREAD(8); break;
      case 9: PGK_SPREAD(9); break;
      default: PGK_SPREAD(10); break;
    }
#undef PGK_SPREAD
  }
  HIP_CHECK(hipGetLastError());
  pgk_prof_end(c, PGK_PROF_SPREAD);
  pgk_prof_begin(c, PGK_PROF_FFT);
  const void *tw = (PREC == PGK_PREC_FP64) ? (const void *)c->d_tw : (const void *)c->d_twf;
  size_t csz = sizeof(C);
  const C *twc = reinterpret_cast<const C *>(tw);
  FftPlan p1 = make_plan(K1), p2 = make_plan(K2), p3 = make_plan(K3);
  int nrow = K2 * K3;
  ConvArgs ca;
  ca.bsp = c->d_bsp;
  for (int d = 0; d < 3; d++) {
    ca.recip[d] = 1.0 / c->box[d];
    ca.K[d] = c->nfft[d];
  }
  double pi = 3.14159265358979323846;
  ca.fac = pi * pi / (c->alpha * c->alpha);
  ca.rk2 = c->rkcut * c->rkcut;
  ca.pivol = pi * c->box[0] * c->box[1] * c->box[2];
  ca.rpivol = 1.0 / ca.pivol;
  ca.netq = c->netchrgsq;
  ca.H = H;
  ca.ifnetq = c->ifnetq;
  // register FFT where the line length has one (32, 64, 128, 256), batched
  // Stockham in LDS otherwise; each axis chooses on its own
  const void *tw1 = twc + c->tw_off[0], *tw2 = twc + c->tw_off[1], *tw3 = twc + c->tw_off[2];
  // small grids: x and y of a z plane in one block
  const size_t plane_lds = fft_plane_bytes(K1, K2, csz);
  const int nlx = lines_for(K1, csz, 16);
  const int bx = (nrow + 2 * nlx - 1) / (2 * nlx);
  bool rx = false;
  if (plane_lds) {
    const int lr = pgk_lds_limit((const void *)k_fft_xy_fwd<PREC>, plane_lds);
    if (lr) return lr;
    k_fft_xy_fwd<PREC><<<K3, 256, plane_lds, st>>>(c->d_grid, c->d_cgrid, K1, K2, H, p1, tw1,
                                                   p2, tw2);
  } else {
    rx = with_reg_fft(K1, [&](auto f) {
      using F = decltype(f);
      const int nb = ((nrow + 1) / 2 + F::LPW - 1) / F::LPW;
      k_fft_x_fwd_reg<PREC, F><<<nb, 64, 0, st>>>(c->d_grid, c->d_cgrid, nrow, tw1, H, K2);
    });
    if (!rx)
      k_fft_x_fwd<PREC><<<bx, 256, (2 * nlx + 1) * K1 * csz, st>>>(c->d_grid, c->d_cgrid, nrow,
                                                                   nlx, p1, tw1, H, K2);
  }
  HIP_CHECK(hipGetLastError());
  const int twy = lines_for(K2, csz, 16);
  const int by = K3 * ((H + twy - 1) / twy);
  auto fft_y = [&](auto sgn) {
    constexpr int SGN = decltype(sgn)::value;
    const bool ry = with_reg_fft(K2, [&](auto f) {
      using F = decltype(f);
      const int nb = (H * K3 + F::LPW - 1) / F::LPW;
      k_fft_y_reg<PREC, F, SGN><<<nb, 64, 0, st>>>(c->d_cgrid, H, K3, tw2);
    });
    if (!ry)
      k_fft_y<PREC><<<by, 256, fft_lds_bytes(twy, K2, csz), st>>>(c->d_cgrid, H, K2, twy, p2,
                                                                  tw2, SGN);
  };
  if (!plane_lds) fft_y(std::integral_constant<int, -1>{});
  HIP_CHECK(hipGetLastError());
  *off = c->red_used;
  double *red = c->d_red + (size_t)c->red_used * PGK_REC;
  const bool rz = with_reg_fft(K3, [&](auto f) {
    using F = decltype(f);
    const int nb = (H * K2 + F::LPW - 1) / F::LPW;
    *nrec = nb;
    k_fft_z_conv_reg<PREC, F><<<nb, 64, 0, st>>>(c->d_cgrid, tw3, ca, red);
  });
  if (!rz) {
    const int twz = lines_for(K3, csz, 16);
    const int bz = K2 * ((H + twz - 1) / twz);
    *nrec = bz;
    k_fft_z_conv<PREC><<<bz, 256, fft_lds_bytes(twz, K3, csz) + (twz + K3) * sizeof(double),
                         st>>>(c->d_cgrid, twz, p3, tw3, ca, red);
  }
  c->red_used += *nrec;
  HIP_CHECK(hipGetLastError());
  // potential grid back into the real grid buffer (compute type)
  if (plane_lds) {
    const int lr = pgk_lds_limit((const void *)k_fft_yx_bwd<PREC>, plane_lds);
    if (lr) return lr;
    k_fft_yx_bwd<PREC><<<K3, 256, plane_lds, st>>>(c->d_cgrid, c->d_pgrid, K1, K2, H, p1, tw1,
                                                   p2, tw2);
  } else {
    fft_y(std::integral_constant<int, +1>{});
    if (!rx)
      k_fft_x_bwd<PREC><<<bx, 256, (2 * nlx + 1) * K1 * csz, st>>>(c->d_cgrid, c->d_pgrid, nrow,
                                                                   nlx, p1, tw1, H, K2);
    else
      with_reg_fft(K1, [&](auto f) {
        using F = decltype(f);
        const int nb = ((nrow + 1) / 2 + F::LPW - 1) / F::LPW;
        k_fft_x_bwd_reg<PREC, F><<<nb, 64, 0, st>>>(c->d_cgrid, c->d_pgrid, nrow, tw1, H, K2);
      });
  }
  HIP_CHECK(hipGetLastError());
  pgk_prof_end(c, PGK_PROF_FFT);
  pgk_prof_begin(c, PGK_PROF_GATHER);
  A *fat = reinterpret_cast<A *>(c->d_fat), *fatc = reinterpret_cast<A *>(c->d_fat_c);
  {
    const int gb = pgk_grid_blocks((int64_t)nitem * 4, 256);   // four lanes per slot
    const int gbx = xcd_grid(gb);
#define PGK_GATHER(SP, P)                                                                 \
  k_gather_t<PREC, SP, P><<<gbx, 256, 0, st>>>(nitem, ord, c->d_pos, c->d_q, g, c->d_pgrid, \
                                              fat, fatc)
    const int32_t *ord = L.order;
    if (split) {
      switch (c->order) {
        case 3: PGK_GATHER(true, 3); break;
        case 4: PGK_GATHER(true, 4); break;
        case 5: PGK_GATHER(true, 5); break;
        case 6: PGK_GATHER(true, 6); break;
        case 7: PGK_GATHER(true, 7); break;
        case 8: PGK_GATHER(true, 8); break;
        case 9: PGK_GATHER(true, 9); break;
        default: PGK_GATHER(true, 10); break;
      }
    } else {
      switch (c->order) {
        case 3: PGK_GATHER(false, 3); break;
        case 4: PGK_GATHER(false, 4); break;
        case 5: PGK_GATHER(false, 5); break;
        case 6: PGK_GATHER(false, 6); break;
        case 7: PGK_GATHER(false, 7); break;
        case 8: PGK_GATHER(false, 8); break;
        case 9: PGK_GATHER(false, 9); break;
        default: PGK_GATHER(false, 10); break;
      }
    }
#undef PGK_GATHER
  }
  HIP_CHECK(hipGetLastError());
  pgk_prof_end(c, PGK_PROF_GATHER);
  return PGK_OK;
}

}  // namespace

// Queue the per-slot preparation (B-spline weights, tile ranks and, with
// pack, the pair kernel's packed positions) ahead of the force kernels.
int pgk_pme_prep(pgk_ctx *c, bool pack) {
  switch (c->precision) {
    case PGK_PREC_FP64: return pack ? launch_prep<PGK_PREC_FP64, true>(c) : launch_prep<PGK_PREC_FP64, false>(c);
    case PGK_PREC_MIXED: return pack ? launch_prep<PGK_PREC_MIXED, true>(c) : launch_prep<PGK_PREC_MIXED, false>(c);
    default: return pack ? launch_prep<PGK_PREC_FP32, true>(c) : launch_prep<PGK_PREC_FP32, false>(c);
  }
}

int pgk_pme_prepare(pgk_ctx *c) {
  for (int d = 0; d < 3; d++) {
    FftPlan p = make_plan(c->nfft[d]);
    if (p.nrad < 0) {
      pgk_set_error("SPME grid sizes must be products of 2, 3 and 5");
      return PGK_EINVAL;
    }
    if (p.nrad > 24) {
      pgk_set_error("SPME grid size too large");
      return PGK_EINVAL;
    }
  }
  if (c->order < 3 || c->order > PGK_MAXORD) {
    pgk_set_error("pme_order must be in 3..10");
    return PGK_EINVAL;
  }
  for (int d = 0; d < 3; d++)
    if (c->nfft[d] < c->order) {
      // fill_charge_grid wraps a stencil index at most once (fft_pme.F90:393-411)
      pgk_set_error("SPME grid must have at least pme_order points per axis");
      return PGK_EINVAL;
    }
  return PGK_OK;
}

int pgk_pme_launch(pgk_ctx *c, bool split, int *off, int *nrec) {
  switch (c->precision) {
    case PGK_PREC_FP64: return launch_pme<PGK_PREC_FP64>(c, split, off, nrec);
    case PGK_PREC_MIXED: return launch_pme<PGK_PREC_MIXED>(c, split, off, nrec);
    default: return launch_pme<PGK_PREC_FP32>(c, split, off, nrec);
  }
}

// per-atom reciprocal energies of the last call's potential grid into d_e
template <int PREC>
static int launch_phi_atoms(pgk_ctx *c, double *d_e) {
  PmeGeom g;
  for (int d = 0; d < 3; d++) {
    g.K[d] = c->nfft[d];
    g.recip[d] = 1.0 / c->box[d];
  }
  g.H = c->nfft[0] / 2 + 1;
  g.order = c->order;
  const int n = c->natom, nb = pgk_grid_blocks(n, 256);
#define PGK_PHI(P) k_phi_atoms<PREC, P><<<nb, 256, 0, c->stream>>>(n, c->d_pos, c->d_q, g, \
                                                                  c->d_pgrid, d_e)
  switch (c->order) {
    case 3: PGK_PHI(3); break;
    case 4: PGK_PHI(4); break;
    case 5: PGK_PHI(5); break;
    case 6: PGK_PHI(6); break;
    case 7: PGK_PHI(7); break;
    case 8: PGK_PHI(8); break;
    case 9: PGK_PHI(9); break;
    default: PGK_PHI(10); break;
  }
#undef PGK_PHI
  HIP_CHECK(hipGetLastError());
  return PGK_OK;
}

int pgk_pme_atom_energy(pgk_ctx *c, double *d_e) {
  switch (c->precision) {
    case PGK_PREC_FP64: return launch_phi_atoms<PGK_PREC_FP64>(c, d_e);
    case PGK_PREC_MIXED: return launch_phi_atoms<PGK_PREC_MIXED>(c, d_e);
    default: return launch_phi_atoms<PGK_PREC_FP32>(c, d_e);
  }
}

// Per-atom reciprocal virial tensors of the charge grid of the last
// reciprocal evaluation (still in d_grid): for each of the six components
// the grid is transformed, multiplied by G_ab, transformed back into
// d_vgrid and interpolated at the atoms as the energies are,
// W_i^ab = q_i / 2 sum theta_i phi_ab; d_w receives (3,3,natom).
template <int PREC>
static int launch_atom_virial(pgk_ctx *c, double *d_w) {
  using R = typename Prec<PREC>::real;
  using C = typename C2T<R>::type;
  hipStream_t st = c->stream;
  const int K1 = c->nfft[0], K2 = c->nfft[1], K3 = c->nfft[2];
  const int H = K1 / 2 + 1, n = c->natom;
  const size_t G = (size_t)K1 * K2 * K3;
  int rc;
  if ((rc = grow_dev(&c->d_vgrid, &c->vgrid_bytes, G * sizeof(R))) ||
      (rc = grow_dev((void **)&c->d_vir6, &c->vir6_bytes, sizeof(double) * 6 * (size_t)n)))
    return rc;
  const void *tw = (PREC == PGK_PREC_FP64) ? (const void *)c->d_tw : (const void *)c->d_twf;
  const size_t csz = sizeof(C);
  const C *twc = reinterpret_cast<const C *>(tw);
  FftPlan p1 = make_plan(K1), p2 = make_plan(K2), p3 = make_plan(K3);
  const int nrow = K2 * K3;
  ConvArgs ca;
  ca.bsp = c->d_bsp;
  for (int d = 0; d < 3; d++) {
    ca.recip[d] = 1.0 / c->box[d];
    ca.K[d] = c->nfft[d];
  }
  const double pi = 3.14159265358979323846;
  ca.fac = pi * pi / (c->alpha * c->alpha);
  ca.rk2 = c->rkcut * c->rkcut;
  ca.pivol = pi * c->box[0] * c->box[1] * c->box[2];
  ca.rpivol = 1.0 / ca.pivol;
  ca.netq = c->netchrgsq;
  ca.H = H;
  ca.ifnetq = 0;
  const void *tw1 = twc + c->tw_off[0], *tw2 = twc + c->tw_off[1], *tw3 = twc + c->tw_off[2];
  // small grids: x and y of a z plane in one block
  const size_t plane_lds = fft_plane_bytes(K1, K2, csz);
  const int nlx = lines_for(K1, csz, 16);
  const int bx = (nrow + 2 * nlx - 1) / (2 * nlx);
  const int twy = lines_for(K2, csz, 16);
  const int by = K3 * ((H + twy - 1) / twy);
  const int twz = lines_for(K3, csz, 16);
  const int bz = K2 * ((H + twz - 1) / twz);
  PmeGeom g;
  for (int d = 0; d < 3; d++) {
    g.K[d] = c->nfft[d];
    g.recip[d] = 1.0 / c->box[d];
  }
  g.H = H;
  g.order = c->order;
  for (int comp = 0; comp < 6; comp++) {
    const bool rx = with_reg_fft(K1, [&](auto f) {
      using F = decltype(f);
      const int nb = ((nrow + 1) / 2 + F::LPW - 1) / F::LPW;
      k_fft_x_fwd_reg<PREC, F><<<nb, 64, 0, st>>>(c->d_grid, c->d_cgrid, nrow, tw1, H, K2);
    });
    if (!rx)
      k_fft_x_fwd<PREC><<<bx, 256, (2 * nlx + 1) * K1 * csz, st>>>(c->d_grid, c->d_cgrid, nrow,
                                                                   nlx, p1, tw1, H, K2);
    auto fft_y = [&](auto sgn) {
      constexpr int SGN = decltype(sgn)::value;
      const bool ry = with_reg_fft(K2, [&](auto f) {
        using F = decltype(f);
        const int nb = (H * K3 + F::LPW - 1) / F::LPW;
        k_fft_y_reg<PREC, F, SGN><<<nb, 64, 0, st>>>(c->d_cgrid, H, K3, tw2);
      });
      if (!ry)
        k_fft_y<PREC><<<by, 256, fft_lds_bytes(twy, K2, csz), st>>>(c->d_cgrid, H, K2, twy, p2,
                                                                    tw2, SGN);
    };
    fft_y(std::integral_constant<int, -1>{});
    k_fft_z_vir<PREC><<<bz, 256, fft_lds_bytes(twz, K3, csz) + (twz + K3) * sizeof(double),
                        st>>>(c->d_cgrid, twz, p3, tw3, ca, comp);
    fft_y(std::integral_constant<int, +1>{});
    if (!rx)
      k_fft_x_bwd<PREC><<<bx, 256, (2 * nlx + 1) * K1 * csz, st>>>(c->d_cgrid, c->d_vgrid, nrow,
                                                                   nlx, p1, tw1, H, K2);
    else
      with_reg_fft(K1, [&](auto f) {
        using F = decltype(f);
        const int nb = ((nrow + 1) / 2 + F::LPW - 1) / F::LPW;
        k_fft_x_bwd_reg<PREC, F><<<nb, 64, 0, st>>>(c->d_cgrid, c->d_vgrid, nrow, tw1, H, K2);
      });
    HIP_CHECK(hipGetLastError());
    double *out = c->d_vir6 + (size_t)comp * n;
    const int nb = pgk_grid_blocks(n, 256);
#define PGK_PHI(P) k_phi_atoms<PREC, P><<<nb, 256, 0, st>>>(n, c->d_pos, c->d_q, g, c->d_vgrid, out)
    switch (c->order) {
      case 3: PGK_PHI(3); break;
      case 4: PGK_PHI(4); break;
      case 5: PGK_PHI(5); break;
      case 6: PGK_PHI(6); break;
      case 7: PGK_PHI(7); break;
      case 8: PGK_PHI(8); break;
      case 9: PGK_PHI(9); break;
      default: PGK_PHI(10); break;
    }
#undef PGK_PHI
    HIP_CHECK(hipGetLastError());
  }
  k_vir_assemble<<<pgk_grid_blocks(n, 256), 256, 0, st>>>(n, c->d_vir6, d_w);
  HIP_CHECK(hipGetLastError());
  return PGK_OK;
}

int pgk_pme_atom_virial(pgk_ctx *c, double *d_w) {
  switch (c->precision) {
    case PGK_PREC_FP64: return launch_atom_virial<PGK_PREC_FP64>(c, d_w);
    case PGK_PREC_MIXED: return launch_atom_virial<PGK_PREC_MIXED>(c, d_w);
    default: return launch_atom_virial<PGK_PREC_FP32>(c, d_w);
  }
}
