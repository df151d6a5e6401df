// dp_device.h -- device-side data layout and helpers shared by the hipcc-built
// pipeline kernels (dp_kernels.hip) and the hiprtc-built pair kernels
// (dp_pair.hip).  Plain C structs only, so that the host library, hipcc and
// hiprtc all see the same layout.
#pragma once

#ifdef __HIPCC_RTC__
typedef unsigned int u32;
typedef int i32;
typedef unsigned long long u64;
typedef long long i64;
#else
#include <hip/hip_runtime.h>
#include <stdint.h>
typedef uint32_t u32;
typedef int32_t i32;
typedef uint64_t u64;
typedef int64_t i64;
#endif

#define DP_DMAX 8          // largest spatial dimension supported
#define DP_MAXTYPES 8      // largest number of particle types (sigma/eps tables)

// One independent system ("replica").  Particles of replica r occupy slots
// [p0, p0+n) of every per-particle array; its cells occupy [c0, c0+ncell).
struct DpRep {
  i64 p0, n;            // particle slot range
  i64 c0, ncell;        // global cell range
  i32 nc[DP_DMAX];      // cells per dimension (>= 3, or 1 = collapsed)
  double L[DP_DMAX];    // box lengths
  double halfL[DP_DMAX];// L/2 (reference: box_vectors[d] / 2)
  double w[DP_DMAX];    // cell width L/nc
  double inv_w[DP_DMAX];// nc/L
  float wf_hi[DP_DMAX]; // w as a float pair (hi + lo) for mixed precision
  float wf_lo[DP_DMAX];
  float Lf[DP_DMAX];    // box lengths as floats (collapsed dims only)
  float invLf[DP_DMAX];
  double nl_thresh;     // s < nl_thresh  <=>  sqrt(s) < cutoff (exact, FP64)
  double skin_half_sq;  // (0.5 * skin)^2
  double dt, T, tau;    // integrator parameters
  i32 collapsed_mask;   // bit d set if nc[d] == 1
  i32 pad_;
};

// Device control block (one per context).
struct DpCtl {
  u64 step_counter;     // steps taken by the device (RNG counter)
  u64 step_index;       // counter value used by the current step
  i32 rebuild;          // 1 if the current step rebuilds the neighbour list
  i32 force_rebuild;    // host request
  i32 overflow;         // some row needed more than K entries
  i32 max_count;        // largest row length seen at the last emit
  i32 nonfinite;        // a non-finite position was seen while binning
  i32 pad_;
};

// Pair-interaction tables (per context; indexed [ti * ntypes + tj]).
struct DpPairTab {
  double sigma[DP_MAXTYPES * DP_MAXTYPES];
  double eps[DP_MAXTYPES * DP_MAXTYPES];
  double cut_thresh[DP_MAXTYPES * DP_MAXTYPES];   // s > t  <=>  sqrt(s)/sigma > r_cut
  double inv_sigma[DP_MAXTYPES * DP_MAXTYPES];
  double inv_sigma2[DP_MAXTYPES * DP_MAXTYPES];
  float sigma_f[DP_MAXTYPES * DP_MAXTYPES];
  float eps_f[DP_MAXTYPES * DP_MAXTYPES];
  float eps_over_sigma_f[DP_MAXTYPES * DP_MAXTYPES];
  float inv_sigma_f[DP_MAXTYPES * DP_MAXTYPES];
  float inv_sigma2_f[DP_MAXTYPES * DP_MAXTYPES];
  float thresh_f[DP_MAXTYPES * DP_MAXTYPES];      // r^2 cut in FP32 (centre of guard band)
  float guard_lo_f[DP_MAXTYPES * DP_MAXTYPES];    // below: certainly inside
  float guard_hi_f[DP_MAXTYPES * DP_MAXTYPES];    // above: certainly outside
  double r_cut;
  double u_rc;          // u(r_cut) used for the energy shift
};

// Arguments of the hiprtc-built pair / step kernels (passed by value).
struct DpPairArgs {
  // inputs (current buffers)
  const double* x;        // [N*D] FP64 positions, slot order
  const float* xs;        // [N*RS] mixed records (cell-relative FP32 + type)
  const u32* nl;          // [K*Ns] neighbour list, column major
  const i32* nlcnt;       // [N]
  const i32* type;        // [N]
  const i32* rep;         // [N]
  const DpRep* reps;      // [R]
  const DpPairTab* tab;
  i64 N, Ns;
  // observable outputs (nullable)
  double* F;              // [N*D]
  double* E;              // [N]
  double* W;              // [N]
  double* LAP;            // [N]
  // integrator state (fused step kernels)
  double* x_out;          // [N*D]
  float* xs_out;          // [N*RS]
  double* v;              // [N*D]
  double* beta;           // [N*D]
  i32* img;               // [N*D]
  const double* mass;     // [N]
  const double* xref;     // [N*D]
  const i32* cellpack;    // [N]
  const i32* orig;        // [N]
  const double* xi;       // injected normals [N*D] in original order, or null
  u64 seed;
  DpCtl* ctl;
  u64* dispmax;           // [R] bit patterns of max displacement^2
  i32 wrap;               // 1: wrap positions into the box after the update
  i32 pad_;
};

// Neighbour-list entry: slot index << (2*D) | 2-bit stencil code per dim
// (code = s + 1 with s in {-1,0,1}; j sits in cell c_i + s).
#define DP_NL_SHIFT(D) (2 * (D))

#ifndef DP_NO_DEVICE_HELPERS

// ------------------------------------------------------------------------
// Philox4x32-10 (Salmon, Moraes, Dror, Shaw, SC'11), counter-based RNG.
// ------------------------------------------------------------------------
struct DpU4 { u32 x, y, z, w; };

__device__ __forceinline__ void dp_mulhilo(u32 a, u32 b, u32& hi, u32& lo) {
  u64 p = (u64)a * (u64)b;
  hi = (u32)(p >> 32);
  lo = (u32)p;
}

__device__ __forceinline__ DpU4 dp_philox4x32_10(DpU4 c, u32 k0, u32 k1) {
  const u32 M0 = 0xD2511F53u, M1 = 0xCD9E8D57u;
  const u32 W0 = 0x9E3779B9u, W1 = 0xBB67AE85u;
#pragma unroll
  for (int r = 0; r < 10; ++r) {
    u32 hi0, lo0, hi1, lo1;
    dp_mulhilo(M0, c.x, hi0, lo0);
    dp_mulhilo(M1, c.z, hi1, lo1);
    DpU4 n;
    n.x = hi1 ^ c.y ^ k0;
    n.y = lo1;
    n.z = hi0 ^ c.w ^ k1;
    n.w = lo0;
    c = n;
    k0 += W0;
    k1 += W1;
  }
  return c;
}

// Two standard normals from 128 random bits (Box-Muller on two 64-bit
// uniforms in (0,1)).
__device__ __forceinline__ void dp_box_muller(DpU4 b, double& z0, double& z1) {
  const double two_m64 = 5.421010862427522170037e-20;  // 2^-64
  u64 a = ((u64)b.x << 32) | b.y;
  u64 c = ((u64)b.z << 32) | b.w;
  // (a >> 11) gives 53 bits; +0.5 keeps u strictly inside (0,1)
  double u1 = ((double)(a >> 11) + 0.5) * 1.1102230246251565404e-16;  // 2^-53
  double u2 = ((double)(c >> 11) + 0.5) * 1.1102230246251565404e-16;
  (void)two_m64;
  double rad = sqrt(-2.0 * log(u1));
  double th = 6.283185307179586476925 * u2;
  double s, co;
  sincos(th, &s, &co);
  z0 = rad * co;
  z1 = rad * s;
}

// Standard normal number `d` (0 <= d < D) for particle `pid` at step `step`.
__device__ __forceinline__ double dp_normal(u64 seed, u64 step, u32 pid, int d) {
  DpU4 c;
  c.x = pid;
  c.y = (u32)(d >> 1);
  c.z = (u32)step;
  c.w = (u32)(step >> 32);
  DpU4 r = dp_philox4x32_10(c, (u32)seed, (u32)(seed >> 32));
  double z0, z1;
  dp_box_muller(r, z0, z1);
  return (d & 1) ? z1 : z0;
}

// ------------------------------------------------------------------------
// Exact FP64 helpers (no contraction: reference arithmetic is un-fused).
// ------------------------------------------------------------------------
__device__ __forceinline__ double dp_mul(double a, double b) { return __dmul_rn(a, b); }
__device__ __forceinline__ double dp_add(double a, double b) { return __dadd_rn(a, b); }
__device__ __forceinline__ double dp_sub(double a, double b) { return __dsub_rn(a, b); }

// positions.py:6-17 ordering: "> L/2" tested first.
__device__ __forceinline__ double dp_minimage_gt(double d, double L, double halfL) {
  if (d > halfL) return dp_sub(d, L);
  if (d < -halfL) return dp_add(d, L);
  return d;
}
// neighbor_list.py:27-31 ordering: "< -L/2" tested first (same result).
__device__ __forceinline__ double dp_minimage_lt(double d, double L, double halfL) {
  if (d < -halfL) return dp_add(d, L);
  if (d > halfL) return dp_sub(d, L);
  return d;
}

#endif  // DP_NO_DEVICE_HELPERS
