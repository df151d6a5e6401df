// dp_pair.hip -- pair force / energy / virial / Laplacian kernels and the
// fused force+integrate step kernels.  Compiled at run time with hiprtc for
// gfx950, once per (dimension, precision mode, number of types, potential);
// the host prepends
//   #define DP_D <dim>         spatial dimension
//   #define DP_MODE <0|1>      0 = FP64 parity mode, 1 = mixed FP32/FP64
//   #define DP_NTYPES <n>      number of particle types
//   #define DP_EVEN <0|1>      1 if f(x)/x and u(x) are functions of x^2
//   #define DP_RS <n>          floats per mixed-precision position record
//   #define DP_ACC <0|1>       mixed accumulator: 0 = FP64, 1 = FP32 Kahan
// followed by the SymPy-generated device functions
//   template<class T> T dp_pot_u(T x);   u(x)
//   template<class T> T dp_pot_f(T x);   f(x) = -u'(x)
//   template<class T> T dp_pot_c(T x);   u''(x)
//   template<class T> T dp_pot_gs(T s);  f(x)/x with s = x^2   (DP_EVEN)
//   template<class T> T dp_pot_us(T s);  u(x) with s = x^2      (DP_EVEN)
//
// Reference semantics (dompap, potential.py:163-188 forces, :69-100 energy,
// :253-279 virial, :327-364 Laplacian, integrator.py:5-40 integrators):
//   x = r / sigma; f_cut(x) = 0 if x > r_cut else f(x) (potential.py:64);
//   F_n += eps * f(x) * d / r  (no 1/sigma factor: potential.py:185);
//   energy term eps * (u(x) - u(r_cut)); virial term eps * f(x) * r;
//   Laplacian term eps * u''(x) + (D - 1) * (-eps * f(x)) / r.
#include "dp_device.h"

#ifndef DP_D
#error "DP_D must be defined"
#endif

#define DP_SB (2 * DP_D)
#define DP_CODE_MASK ((1u << DP_SB) - 1u)
#define DP_PACK_BITS ((DP_D == 1) ? 31 : 32 / DP_D)

__device__ __forceinline__ int dp_tt(int ti, int tj) { return ti * DP_NTYPES + tj; }

// Sum of pair terms for particle i; writes forces to f[] and the scalar sums.
template <bool WANT_E, bool WANT_W, bool WANT_L>
__device__ __forceinline__ void dp_pair_sum(const DpPairArgs& a, i64 i, const DpRep& R,
                                            double f[DP_D], double& e, double& w, double& lap) {
  const DpPairTab& T = *a.tab;
  const int cnt = a.nlcnt[i];
  const u32* nlp = a.nl + i;
  const i64 Ns = a.Ns;
#pragma unroll
  for (int d = 0; d < DP_D; ++d) f[d] = 0.0;
  e = 0.0; w = 0.0; lap = 0.0;

#if DP_MODE == 0
  // ---------------- FP64 parity mode ----------------
  double xi_[DP_D];
#pragma unroll
  for (int d = 0; d < DP_D; ++d) xi_[d] = a.x[i * DP_D + d];
  const int ti = (DP_NTYPES > 1) ? a.type[i] : 0;
  for (int k = 0; k < cnt; ++k) {
    const u32 ent = nlp[(i64)k * Ns];
    const i64 j = (i64)(ent >> DP_SB);
    double dd[DP_D];
    double s = 0.0;
#pragma unroll
    for (int d = 0; d < DP_D; ++d) {
      double t = dp_sub(xi_[d], a.x[j * DP_D + d]);
      t = dp_minimage_gt(t, R.L[d], R.halfL[d]);
      dd[d] = t;
      s = (d == 0) ? dp_mul(t, t) : dp_add(s, dp_mul(t, t));
    }
    const int tt = dp_tt(ti, (DP_NTYPES > 1) ? a.type[j] : 0);
    if (s > T.cut_thresh[tt]) continue;          // x > r_cut  <=>  s > threshold
    const double r = sqrt(s);
    const double xr = r / T.sigma[tt];
    const double eps = T.eps[tt];
    const double sf = eps * dp_pot_f<double>(xr);
#pragma unroll
    for (int d = 0; d < DP_D; ++d) f[d] = f[d] + sf * (dd[d] / r);
    if (WANT_E) e += eps * (dp_pot_u<double>(xr) - T.u_rc);
    if (WANT_W) w += sf * r;
    if (WANT_L) lap += eps * dp_pot_c<double>(xr) + (double)(DP_D - 1) * (-sf) / r;
  }
#else
  // ---------------- mixed FP32 pair / FP64 accumulate ----------------
  float xi_[DP_RS];
#pragma unroll
  for (int q = 0; q < DP_RS; ++q) xi_[q] = a.xs[i * DP_RS + q];
  const int ti = (DP_NTYPES > 1) ? __float_as_int(xi_[DP_D]) : 0;
  float whi[DP_D], wlo[DP_D];
#pragma unroll
  for (int d = 0; d < DP_D; ++d) { whi[d] = R.wf_hi[d]; wlo[d] = R.wf_lo[d]; }
  const int collapsed = R.collapsed_mask;
#if DP_ACC == 1
  float fk[DP_D], ck[DP_D];
#pragma unroll
  for (int d = 0; d < DP_D; ++d) { fk[d] = 0.f; ck[d] = 0.f; }
#endif
  for (int k = 0; k < cnt; ++k) {
    const u32 ent = nlp[(i64)k * Ns];
    const i64 j = (i64)(ent >> DP_SB);
    float xj[DP_RS];
#if DP_RS == 2
    { float2 q = *reinterpret_cast<const float2*>(a.xs + j * 2); xj[0] = q.x; xj[1] = q.y; }
#else
#pragma unroll
    for (int q4 = 0; q4 < DP_RS; q4 += 4) {
      float4 q = *reinterpret_cast<const float4*>(a.xs + j * DP_RS + q4);
      xj[q4] = q.x; xj[q4 + 1] = q.y; xj[q4 + 2] = q.z; xj[q4 + 3] = q.w;
    }
#endif
    float dd[DP_D];
    float s2 = 0.f;
#pragma unroll
    for (int d = 0; d < DP_D; ++d) {
      const float sd = (float)((int)((ent >> (2 * d)) & 3u) - 1);
      float t = xi_[d] - xj[d];
      t = fmaf(-sd, whi[d], t);
      t = fmaf(-sd, wlo[d], t);
      if (collapsed & (1 << d)) t -= R.Lf[d] * rintf(t * R.invLf[d]);
      dd[d] = t;
      s2 = fmaf(t, t, s2);
    }
    const int tt = dp_tt(ti, (DP_NTYPES > 1) ? __float_as_int(xj[DP_D]) : 0);
    if (s2 >= T.guard_hi_f[tt]) continue;       // certainly beyond the cut
    if (s2 > T.guard_lo_f[tt]) {
      // Too close to the cut to decide in FP32: decide exactly in FP64 with
      // the reference's operation order (rare).
      double s = 0.0;
#pragma unroll
      for (int d = 0; d < DP_D; ++d) {
        double t = dp_sub(a.x[i * DP_D + d], a.x[j * DP_D + d]);
        t = dp_minimage_gt(t, R.L[d], R.halfL[d]);
        s = (d == 0) ? dp_mul(t, t) : dp_add(s, dp_mul(t, t));
      }
      if (s > T.cut_thresh[tt]) continue;
    }
    const float sx = s2 * T.inv_sigma2_f[tt];
#if DP_EVEN
    const float g = dp_pot_gs<float>(sx);               // f(x)/x
#else
    const float xr = sqrtf(sx);
    const float g = dp_pot_f<float>(xr) / xr;
#endif
    const float fs = T.eps_over_sigma_f[tt] * g;       // eps f(x) / r
#if DP_ACC == 1
#pragma unroll
    for (int d = 0; d < DP_D; ++d) {
      const float y = fs * dd[d] - ck[d];
      const float t = fk[d] + y;
      ck[d] = (t - fk[d]) - y;
      fk[d] = t;
    }
#else
#pragma unroll
    for (int d = 0; d < DP_D; ++d) f[d] += (double)(fs * dd[d]);
#endif
    if (WANT_E || WANT_W || WANT_L) {
      const float eps = T.eps_f[tt];
      if (WANT_E) {
#if DP_EVEN
        e += (double)(eps * dp_pot_us<float>(sx)) - (double)eps * T.u_rc;
#else
        e += (double)(eps * dp_pot_u<float>(xr)) - (double)eps * T.u_rc;
#endif
      }
      if (WANT_W) w += (double)(eps * T.sigma_f[tt] * g * sx);   // eps f(x) r
      if (WANT_L) {
        const float r = sqrtf(s2);
#if DP_EVEN
        const float xr2 = sqrtf(sx);
        lap += (double)(eps * dp_pot_c<float>(xr2)) - (double)(DP_D - 1) * (double)(fs);
#else
        lap += (double)(eps * dp_pot_c<float>(xr)) - (double)(DP_D - 1) * (double)(fs);
#endif
        (void)r;
      }
    }
  }
#if DP_ACC == 1
#pragma unroll
  for (int d = 0; d < DP_D; ++d) f[d] = (double)fk[d] + (double)(-ck[d]);
#endif
#endif
}

// ---------------------------------------------------------------------------
// Observables: forces (and optionally per-particle energy, virial, Laplacian)
// ---------------------------------------------------------------------------
extern "C" __global__ void __launch_bounds__(256) dp_pair_force(DpPairArgs a) {
  const i64 i = (i64)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= a.N) return;
  const DpRep& R = a.reps[a.rep[i]];
  double f[DP_D], e, w, lap;
  dp_pair_sum<false, false, false>(a, i, R, f, e, w, lap);
#pragma unroll
  for (int d = 0; d < DP_D; ++d) a.F[i * DP_D + d] = f[d];
}

extern "C" __global__ void __launch_bounds__(256) dp_pair_obs(DpPairArgs a) {
  const i64 i = (i64)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= a.N) return;
  const DpRep& R = a.reps[a.rep[i]];
  double f[DP_D], e, w, lap;
  dp_pair_sum<true, true, true>(a, i, R, f, e, w, lap);
  if (a.F) {
#pragma unroll
    for (int d = 0; d < DP_D; ++d) a.F[i * DP_D + d] = f[d];
  }
  if (a.E) a.E[i] = e;
  if (a.W) a.W[i] = w;
  if (a.LAP) a.LAP[i] = lap;
}

// ---------------------------------------------------------------------------
// Fused step: forces of particle i from the current buffers, then the
// integrator update of particle i into the next buffers, the periodic wrap
// (positions.py:90-98), and the skin-displacement maximum
// (neighbor_list.py:183-201) for the next step's rebuild decision.
// ---------------------------------------------------------------------------
__device__ __forceinline__ void dp_dispmax_update(const DpPairArgs& a, bool valid, int r, double s) {
  // Wave-level max when every valid lane of the wavefront belongs to one
  // replica (the common case); per-lane atomics at replica boundaries.
  // Non-negative doubles order like their bit patterns.
  const int lane = threadIdx.x & 63;
  const unsigned long long b = __ballot(valid);
  if (b == 0ull) return;
  const int first = __ffsll((unsigned long long)b) - 1;
  const int rf = __shfl(r, first);
  const bool uniform = __all((!valid) || r == rf);
  if (uniform) {
    double m = valid ? s : 0.0;
#pragma unroll
    for (int off = 32; off >= 1; off >>= 1) m = fmax(m, __shfl_xor(m, off));
    if (lane == first) atomicMax(&a.dispmax[rf], (unsigned long long)__double_as_longlong(m));
  } else if (valid) {
    atomicMax(&a.dispmax[r], (unsigned long long)__double_as_longlong(s));
  }
}

template <int INTEG, bool FORCES_GIVEN>
__device__ __forceinline__ void dp_step_body(const DpPairArgs& a) {
  const i64 i = (i64)blockIdx.x * blockDim.x + threadIdx.x;
  const bool valid = i < a.N;
  const i64 ii = valid ? i : 0;
  const int r = a.rep[ii];
  const DpRep& R = a.reps[r];
  double f[DP_D], e, w, lap;
  if (FORCES_GIVEN) {
#pragma unroll
    for (int d = 0; d < DP_D; ++d) f[d] = a.F[ii * DP_D + d];
  } else if (valid) {
    dp_pair_sum<false, false, false>(a, ii, R, f, e, w, lap);
  }
  double sdisp = 0.0;
  if (valid) {
#pragma clang fp contract(off)
    const double m = a.mass[ii];
    const double dt = R.dt;
    const u64 step = a.ctl->step_index;
    const u32 pid = (u32)a.orig[ii];
    double xn[DP_D];
    if (INTEG == 1) {
      // integrator.py:17-40 (G-JF), numpy's left-to-right evaluation order
      const double alpha = m / R.tau;
      const double bvar = 2.0 * alpha * R.T * dt;
      const double sq = sqrt(bvar);
      const double q = alpha * dt / 2.0 / m;
      const double num = 1.0 - q;
      const double den = 1.0 + q;
      const double ca = num / den;
      const double cb = 1.0 / den;
#pragma unroll
      for (int d = 0; d < DP_D; ++d) {
        const double z = a.xi ? a.xi[(i64)pid * DP_D + d] : dp_normal(a.seed, step, pid, d);
        const double nb = sq * z;
        const double v0 = a.v[ii * DP_D + d];
        const double b0 = a.beta[ii * DP_D + d];
        const double v1 = ca * v0 + cb * dt * f[d] / m + cb * (b0 + nb) / 2.0 / m;
        xn[d] = a.x[ii * DP_D + d] + v1 * dt;
        a.v[ii * DP_D + d] = v1;
        a.beta[ii * DP_D + d] = nb;
      }
    } else {
      // integrator.py:5-14 (leap-frog)
#pragma unroll
      for (int d = 0; d < DP_D; ++d) {
        const double v1 = a.v[ii * DP_D + d] + dt * f[d] / m;
        xn[d] = a.x[ii * DP_D + d] + dt * v1;
        a.v[ii * DP_D + d] = v1;
      }
    }
    // positions.py:90-98 wrap, then neighbor_list.py:183-201 displacement
#pragma unroll
    for (int d = 0; d < DP_D; ++d) {
      if (a.wrap) {
        const double sh = floor(xn[d] / R.L[d]);
        a.img[ii * DP_D + d] += (i32)sh;
        xn[d] = xn[d] - sh * R.L[d];
      }
      a.x_out[ii * DP_D + d] = xn[d];
      double t = dp_sub(xn[d], a.xref[ii * DP_D + d]);
      t = dp_minimage_gt(t, R.L[d], R.halfL[d]);
      sdisp = (d == 0) ? dp_mul(t, t) : dp_add(sdisp, dp_mul(t, t));
    }
#if DP_MODE == 1
    const u32 pack = (u32)a.cellpack[ii];
    float rec[DP_RS];
#pragma unroll
    for (int q = 0; q < DP_RS; ++q) rec[q] = 0.f;
#pragma unroll
    for (int d = 0; d < DP_D; ++d) {
      const u32 c = (pack >> (DP_PACK_BITS * d)) & ((1u << DP_PACK_BITS) - 1u);
      const double cen = ((double)c + 0.5) * R.w[d];
      rec[d] = (float)dp_minimage_gt(xn[d] - cen, R.L[d], R.halfL[d]);
    }
    if (DP_D < DP_RS) rec[DP_D] = a.xs[ii * DP_RS + DP_D];   // type bits
#pragma unroll
    for (int q = 0; q < DP_RS; ++q) a.xs_out[ii * DP_RS + q] = rec[q];
#endif
  }
  dp_dispmax_update(a, valid, r, sdisp);
}

extern "C" __global__ void __launch_bounds__(256) dp_step_gjf(DpPairArgs a) { dp_step_body<1, false>(a); }
extern "C" __global__ void __launch_bounds__(256) dp_step_lf(DpPairArgs a) { dp_step_body<2, false>(a); }
// integrator only, forces supplied in a.F (integrator.py:5-40 as free functions)
extern "C" __global__ void __launch_bounds__(256) dp_integrate_gjf(DpPairArgs a) { dp_step_body<1, true>(a); }
extern "C" __global__ void __launch_bounds__(256) dp_integrate_lf(DpPairArgs a) { dp_step_body<2, true>(a); }
