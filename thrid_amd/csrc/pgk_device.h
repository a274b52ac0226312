// pgk_device.h - device helpers shared by the force kernels.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "pgk_internal.h"

// ---- arithmetic forms ------------------------------------------------------
// FP64:  compute double, accumulate double (hardware FP64 atomics)
// MIXED: compute float, accumulate int64 fixed point (2^-32) - deterministic
// FP32:  compute float, accumulate float
template <int PREC> struct Prec;
template <> struct Prec<PGK_PREC_FP64> {
  using real = double;
  using real4 = double4;
  using acc = double;
};
template <> struct Prec<PGK_PREC_MIXED> {
  using real = float;
  using real4 = float4;
  using acc = unsigned long long;
};
template <> struct Prec<PGK_PREC_FP32> {
  using real = float;
  using real4 = float4;
  using acc = float;
};

__device__ __forceinline__ void acc_add(double *p, double v) { atomicAdd(p, v); }
__device__ __forceinline__ void acc_add(float *p, float v) { atomicAdd(p, v); }
__device__ __forceinline__ void acc_add(float *p, double v) { atomicAdd(p, (float)v); }
__device__ __forceinline__ void acc_add(unsigned long long *p, double v) {
  long long x = __double2ll_rn(v * PGK_FX_SCALE);
  atomicAdd(p, (unsigned long long)x);
}
__device__ __forceinline__ void acc_add(unsigned long long *p, float v) {
  long long x = __double2ll_rn((double)v * PGK_FX_SCALE);
  atomicAdd(p, (unsigned long long)x);
}

__device__ __forceinline__ double acc_load(const double *p) { return *p; }
__device__ __forceinline__ double acc_load(const float *p) { return (double)*p; }
__device__ __forceinline__ double acc_load(const unsigned long long *p) {
  return (double)(long long)(*p) * PGK_FX_INV;
}

// Workgroups are dispatched to the 8 XCDs round-robin (block b -> XCD b % 8)
// and every XCD has its own L2.  This remaps the block index so that each XCD
// walks one contiguous eighth of the work, which for spatially sorted data is
// a compact region.  The grid must be launched with a multiple of 8 blocks
// (xcd_grid); logical blocks past the real work must return.
static inline int xcd_grid(int nblk) { return (nblk + 7) & ~7; }
__device__ __forceinline__ int xcd_block() {
  const int per = (int)(gridDim.x >> 3);
  return (int)(blockIdx.x & 7) * per + (int)(blockIdx.x >> 3);
}

// Fortran dnint/anint: round half away from zero
__device__ __forceinline__ double fnint(double x) { return round(x); }
__device__ __forceinline__ float fnint(float x) { return roundf(x); }

template <class T>
__device__ __forceinline__ T wave_sum(T v) {
#pragma unroll
  for (int o = 32; o >= 1; o >>= 1) v += __shfl_xor(v, o, 64);
  return v;
}

// Block-level deterministic record write: each wave reduces NF values, the
// waves combine in LDS in fixed order, thread 0 writes the record.
template <int NF, int NW>
__device__ __forceinline__ void block_record(double (&v)[NF], double *rec) {
  __shared__ double s[NW][NF];
  int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
#pragma unroll
  for (int f = 0; f < NF; f++) {
    double t = wave_sum(v[f]);
    if (lane == 0) s[wave][f] = t;
  }
  __syncthreads();
  if (threadIdx.x < NF) {
    double t = 0.0;
#pragma unroll
    for (int w = 0; w < NW; w++) t += s[w][threadIdx.x];
    rec[threadIdx.x] = t;
  }
}

// float -> 2^-32 fixed point as hi:lo = floor(v) : (v - floor(v)) 2^32.
// v - floor(v) is exact for |v| >= 1 (Sterbenz) and for 0 <= v < 1; for
// -1 < v < 0 it is rounded once (error <= 2^-24).  A pure function of v, so
// sums of the results are independent of the order of the atomic adds.
__device__ __forceinline__ unsigned long long fx_of(float v) {
  const float fl = floorf(v);
  const float fr = fminf(v - fl, 0x1.fffffep-1f);
  const unsigned lo = (unsigned)(fr * 4294967296.0f);
  const int hi = (int)fl;
  return ((unsigned long long)(unsigned)hi << 32) | (unsigned long long)lo;
}

#define PGK_REC 16   // doubles per reduction record
