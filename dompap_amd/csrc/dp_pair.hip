// dp_pair.hip -- pair force / energy / virial / Laplacian kernels and the
// fused force + integrate step kernels.  Compiled at run time with hiprtc for
// gfx950, once per (dimension, precision mode, number of types, potential,
// tile shape); the host prepends
//   #define DP_D <dim>          spatial dimension (1..4)
//   #define DP_MODE <0|1>       0 = FP64 parity mode, 1 = mixed precision
//   #define DP_NTYPES <n>       number of particle types
//   #define DP_EVEN <0|1>       1 if f(x)/x and u(x) are functions of x^2
//   #define DP_RS <n>           ints per mixed-precision record
//   #define DP_TPP <n>          lanes per particle (1, 2, 4, 8, 16)
//   #define DP_NLW <16|32>      neighbour-entry width in bits
//   #define DP_OB <n>           offset bits of an entry (stencil index above)
//   #define DP_PAIRPREC <32|64> mixed mode: precision of the in-range pair terms
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
//
// Work decomposition: one workgroup per tile of cells at a time (grid-stride
// over tiles, each XCD taking a contiguous chunk of tiles so that the halo
// re-reads stay in its L2).  The halo box's records are staged in LDS with
// coalesced loads; DP_TPP lanes share a particle's neighbour row and reduce
// their partial sums with cross-lane shuffles.
#include "dp_device.h"

#ifndef DP_D
#error "DP_D must be defined"
#endif

#define DP_N3 (DpPow3<DP_D>::value)
#define DP_OMASK ((1u << DP_OB) - 1u)
#define DP_PACK_BITS ((DP_D == 1) ? 31 : 32 / DP_D)
#define DP_GROUPS (DP_TILE_THREADS / DP_TPP)

#if DP_NLW == 16
typedef u16 dp_ent_t;
#else
typedef u32 dp_ent_t;
#endif

__device__ __forceinline__ int dp_tt(int ti, int tj) { return ti * DP_NTYPES + tj; }

template <class T>
__device__ __forceinline__ T dp_group_sum(T v) {
#pragma unroll
  for (int off = DP_TPP / 2; off >= 1; off >>= 1) v += __shfl_xor(v, off);
  return v;
}

// XCD-aware tile order: workgroup b runs on XCD b % 8; the tiles of XCD k are
// the contiguous chunk [k * chunk, (k + 1) * chunk), shared by that XCD's
// workgroups in turn.
__device__ __forceinline__ i64 dp_tile_of(i64 j, i64 ntiles) {
  const i64 G = gridDim.x;
  const i64 chunk = (ntiles + 7) / 8;
  const i64 xcd = blockIdx.x & 7;
  const i64 lane_in_xcd = blockIdx.x >> 3;
  const i64 per_xcd = (G - xcd + 7) / 8;  // workgroups on this XCD
  return xcd * chunk + lane_in_xcd + j * per_xcd;
}

// KIND: 0 forces, 1 observables (forces + per-particle energy / virial /
// Laplacian), 2 fused force + G-JF, 3 fused force + leap-frog.
template <int KIND>
__device__ void dp_tile_run(const DpPairArgs& a) {
  __shared__ DpTileSmem<DP_D> S;
#if DP_MODE == 1
  __shared__ i32 shq[DP_N3 * DP_D];
  extern __shared__ i32 srec[];  // cap * DP_RS
#else
  extern __shared__ double sxd[];  // cap * DP_D (+ cap ints of types)
#endif
  const DpPairTab& T = *a.tab;
  const int tid = threadIdx.x;
  const int group = tid / DP_TPP;
  const int sub = tid % DP_TPP;
  const dp_ent_t* nl = (const dp_ent_t*)a.nl;
  double dmax = 0.0;  // running maximum displacement^2 of this lane's particles
  int cur_rep = -1;
  const i64 chunk = (a.ntiles + 7) / 8;
  const i64 xcd = blockIdx.x & 7;
  const i64 cend = (xcd + 1) * chunk < a.ntiles ? (xcd + 1) * chunk : a.ntiles;
  for (i64 j = 0;; ++j) {
    const i64 t = dp_tile_of(j, a.ntiles);
    if (t >= cend) break;
    dp_tile_setup<DP_D>(S, a.reps, a.R, a.cstart, t, a.cap);
    const int r = S.rep;
    const DpRep& P = a.reps[r];
    if ((KIND == 2 || KIND == 3) && r != cur_rep) {
      // flush the displacement maximum of the previous replica
      if (cur_rep >= 0) {
        double m = dmax;
#pragma unroll
        for (int off = 32; off >= 1; off >>= 1) m = fmax(m, __shfl_xor(m, off));
        if ((tid & 63) == 0) atomicMax(&a.dispmax[cur_rep], (unsigned long long)__double_as_longlong(m));
      }
      dmax = 0.0;
      cur_rep = r;
    }
    const int nh = P.nhalo;
    const bool fits = S.fits != 0;
#if DP_MODE == 1
    for (int q = tid; q < DP_N3 * DP_D; q += blockDim.x) {
      const int st = q / DP_D, d = q % DP_D;
      int sd[DP_D];
      dp_stencil_digits<DP_D>(st, P.order, sd);
      shq[q] = (P.nc[d] == 1) ? 0 : -sd[d] * P.Wq[d];
    }
    if (fits) {
      for (int idx = tid; idx < S.nrec; idx += blockDim.x) {
        const int u = dp_halo_of<DP_D>(S, nh, idx);
        const i64 slot = S.hstart[u] + (idx - S.hbase[u]);
#pragma unroll
        for (int q = 0; q < DP_RS; ++q) srec[idx * DP_RS + q] = a.xq[slot * DP_RS + q];
      }
    }
#else
    i32* stype = (i32*)(sxd + (size_t)a.cap * DP_D);
    if (fits) {
      for (int idx = tid; idx < S.nrec; idx += blockDim.x) {
        const int u = dp_halo_of<DP_D>(S, nh, idx);
        const i64 slot = S.hstart[u] + (idx - S.hbase[u]);
#pragma unroll
        for (int d = 0; d < DP_D; ++d) sxd[idx * DP_D + d] = a.x[slot * DP_D + d];
        if (DP_NTYPES > 1) stype[idx] = a.type[slot];
      }
    }
#endif
    __syncthreads();
    for (int base = 0; base < S.nhomep; base += DP_GROUPS) {
      const int k = base + group;
      const bool active = k < S.nhomep;
      int uh = 0, loc = 0;
      if (active) {
        const int m = dp_home_of<DP_D>(S, P.nhome, k);
        uh = S.home_u[m];
        loc = k - S.home_p[m];
      }
      const i64 slot_i = active ? (i64)S.hstart[uh] + loc : 0;
      const int self = S.hbase[uh] + loc;
      // rows longer than K were truncated at the emit (and reported as overflow)
      const int cnt = active ? min(a.nlcnt[slot_i], a.K) : 0;
      const dp_ent_t* row = nl + slot_i * a.K;
      double f[DP_D];
#pragma unroll
      for (int d = 0; d < DP_D; ++d) f[d] = 0.0;
      double e = 0.0, w = 0.0, lap = 0.0;
#if DP_MODE == 1
      // ---------------- mixed: fixed-point displacements ----------------
      i32 qi[DP_D];
      int ti = 0;
#pragma unroll
      for (int d = 0; d < DP_D; ++d) qi[d] = fits ? srec[self * DP_RS + d] : a.xq[slot_i * DP_RS + d];
      if (DP_NTYPES > 1) ti = fits ? srec[self * DP_RS + DP_D] : a.xq[slot_i * DP_RS + DP_D];
      for (int kk = sub; kk < cnt; kk += DP_TPP) {
        const u32 ent = (u32)row[kk];
        const int st = (int)(ent >> DP_OB);
        const int off = (int)(ent & DP_OMASK);
        const int u = uh + S.st_off[st];
        const int idx = S.hbase[u] + off;
        i32 dq[DP_D];
        float s2 = 0.f;
        int tj = 0;
#pragma unroll
        for (int d = 0; d < DP_D; ++d) {
          const i32 qj = fits ? srec[idx * DP_RS + d] : a.xq[((i64)S.hstart[u] + off) * DP_RS + d];
          i32 dd = (i32)((u32)qi[d] - (u32)qj + (u32)shq[st * DP_D + d]);
          if (P.nc[d] == 1) {
            const i32 h = P.Lq[d] >> 1;
            if (dd > h) dd -= P.Lq[d];
            else if (dd < -h) dd += P.Lq[d];
          }
          dq[d] = dd;
          const float df = (float)dd;
          s2 = fmaf(df, df, s2);
        }
        if (DP_NTYPES > 1) tj = fits ? srec[idx * DP_RS + DP_D] : a.xq[((i64)S.hstart[u] + off) * DP_RS + DP_D];
        const int tt = dp_tt(ti, tj);
        if (s2 >= T.guard_hi_q[tt]) continue;  // certainly beyond the cut
        if (s2 > T.guard_lo_q[tt]) {
          // too close to the cut to decide from the fixed-point values:
          // decide with the reference's FP64 operations (rare)
          const i64 slot_j = (i64)S.hstart[u] + off;
          double s = 0.0;
#pragma unroll
          for (int d = 0; d < DP_D; ++d) {
            const double tq = dp_minimage_gt(dp_sub(a.x[slot_i * DP_D + d], a.x[slot_j * DP_D + d]), P.L[d],
                                             P.halfL[d]);
            s = (d == 0) ? dp_mul(tq, tq) : dp_add(s, dp_mul(tq, tq));
          }
          if (s > T.cut_thresh[tt]) continue;
        }
#if DP_PAIRPREC == 64
        double s64 = 0.0;
#pragma unroll
        for (int d = 0; d < DP_D; ++d) s64 = fma((double)dq[d], (double)dq[d], s64);
        const double sx = s64 * T.qscale2 * T.inv_sigma2[tt];
#if DP_EVEN
        const double g = dp_pot_gs<double>(sx);  // f(x)/x
#else
        const double xr = sqrt(sx);
        const double g = dp_pot_f<double>(xr) * dp_rcp(xr);
#endif
        const double fs = T.eps_over_sigma[tt] * g * T.qscale;  // eps f(x)/r per fixed-point unit
#pragma unroll
        for (int d = 0; d < DP_D; ++d) f[d] = fma(fs, (double)dq[d], f[d]);
        if (KIND == 1) {
          const double eps = T.eps[tt];
#if DP_EVEN
          e += eps * (dp_pot_us<double>(sx) - T.u_rc);
#else
          e += eps * (dp_pot_u<double>(xr) - T.u_rc);
#endif
          w += eps * T.sigma[tt] * g * sx;
          lap += eps * dp_pot_c<double>(sqrt(sx)) - (double)(DP_D - 1) * T.eps_over_sigma[tt] * g;
        }
#else
        const float sxf = s2 * T.inv_sigma2_qf[tt];
#if DP_EVEN
        const float g = dp_pot_gs<float>(sxf);
#else
        const float xrf = sqrtf(sxf);
        const float g = dp_pot_f<float>(xrf) * dp_rcp(xrf);
#endif
        const float fs = T.eps_over_sigma_qf[tt] * g;
#pragma unroll
        for (int d = 0; d < DP_D; ++d) f[d] += (double)(fs * (float)dq[d]);
        if (KIND == 1) {
          const float eps = T.eps_f[tt];
#if DP_EVEN
          e += (double)(eps * dp_pot_us<float>(sxf)) - (double)eps * T.u_rc;
#else
          e += (double)(eps * dp_pot_u<float>(xrf)) - (double)eps * T.u_rc;
#endif
          w += (double)(eps * T.sigma_f[tt] * g * sxf);
          lap += (double)(eps * dp_pot_c<float>(sqrtf(sxf))) -
                 (double)(DP_D - 1) * (double)(T.eps_f[tt] / T.sigma_f[tt] * g);
        }
#endif
      }
#else
      // ---------------- FP64 parity mode ----------------
      double xi_[DP_D];
#pragma unroll
      for (int d = 0; d < DP_D; ++d) xi_[d] = fits ? sxd[self * DP_D + d] : a.x[slot_i * DP_D + d];
      const int ti = (DP_NTYPES > 1) ? (fits ? stype[self] : a.type[slot_i]) : 0;
      for (int kk = sub; kk < cnt; kk += DP_TPP) {
        const u32 ent = (u32)row[kk];
        const int st = (int)(ent >> DP_OB);
        const int off = (int)(ent & DP_OMASK);
        const int u = uh + S.st_off[st];
        const int idx = S.hbase[u] + off;
        const i64 slot_j = (i64)S.hstart[u] + off;
        double dd[DP_D];
        double s = 0.0;
#pragma unroll
        for (int d = 0; d < DP_D; ++d) {
          const double xj = fits ? sxd[idx * DP_D + d] : a.x[slot_j * DP_D + d];
          const double tq = dp_minimage_gt(dp_sub(xi_[d], xj), P.L[d], P.halfL[d]);
          dd[d] = tq;
          s = (d == 0) ? dp_mul(tq, tq) : dp_add(s, dp_mul(tq, tq));
        }
        const int tj = (DP_NTYPES > 1) ? (fits ? stype[idx] : a.type[slot_j]) : 0;
        const int tt = dp_tt(ti, tj);
        if (s > T.cut_thresh[tt]) continue;  // x > r_cut  <=>  s > threshold
        const double rr = sqrt(s);
        const double xr = rr / T.sigma[tt];
        const double eps = T.eps[tt];
        const double sf = eps * dp_pot_f<double>(xr);
#pragma unroll
        for (int d = 0; d < DP_D; ++d) f[d] = f[d] + sf * (dd[d] / rr);
        if (KIND == 1) {
          e += eps * (dp_pot_u<double>(xr) - T.u_rc);
          w += sf * rr;
          lap += eps * dp_pot_c<double>(xr) + (double)(DP_D - 1) * (-sf) / rr;
        }
      }
#endif
      // combine the partial sums of the particle's DP_TPP lanes
#pragma unroll
      for (int d = 0; d < DP_D; ++d) f[d] = dp_group_sum(f[d]);
      if (KIND == 1) {
        e = dp_group_sum(e);
        w = dp_group_sum(w);
        lap = dp_group_sum(lap);
      }
      if (!active || sub != 0) continue;
      if (KIND == 0 || KIND == 1) {
        if (a.F) {
#pragma unroll
          for (int d = 0; d < DP_D; ++d) a.F[slot_i * DP_D + d] = f[d];
        }
        if (KIND == 1) {
          if (a.E) a.E[slot_i] = e;
          if (a.W) a.W[slot_i] = w;
          if (a.LAP) a.LAP[slot_i] = lap;
        }
        continue;
      }
      // fused integrator update of particle slot_i (one lane per particle)
      {
#pragma clang fp contract(off)
        const i64 ii = slot_i;
        const double m = a.mass[ii];
        const double dt = P.dt;
        const u64 step = a.ctl->step_index;
        const u32 pid = (u32)a.orig[ii];
        double xn[DP_D];
        if (KIND == 2) {
          // integrator.py:17-40 (G-JF), numpy's left-to-right evaluation order
          const double alpha = m / P.tau;
          const double bvar = 2.0 * alpha * P.T_ * dt;
          const double sq = sqrt(bvar);
          const double qq = alpha * dt / 2.0 / m;
          const double num = 1.0 - qq;
          const double den = 1.0 + qq;
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
        double sdisp = 0.0;
#pragma unroll
        for (int d = 0; d < DP_D; ++d) {
          if (a.wrap) {
            const double sh = floor(xn[d] / P.L[d]);
            a.img[ii * DP_D + d] += (i32)sh;
            xn[d] = xn[d] - sh * P.L[d];
          }
          a.x_out[ii * DP_D + d] = xn[d];
          double tq = dp_sub(xn[d], a.xref[ii * DP_D + d]);
          tq = dp_minimage_gt(tq, P.L[d], P.halfL[d]);
          sdisp = (d == 0) ? dp_mul(tq, tq) : dp_add(sdisp, dp_mul(tq, tq));
        }
        if (sdisp != sdisp) sdisp = __longlong_as_double(0x7ff0000000000000ll);
        dmax = fmax(dmax, sdisp);
#if DP_MODE == 1
        const u32 pack = (u32)a.cellpack[ii];
#pragma unroll
        for (int d = 0; d < DP_D; ++d) {
          const u32 c = (pack >> (DP_PACK_BITS * d)) & ((1u << DP_PACK_BITS) - 1u);
          const double cen = ((double)c + 0.5) * P.w[d];
          a.xq_out[ii * DP_RS + d] = (i32)rint(dp_minimage_gt(xn[d] - cen, P.L[d], P.halfL[d]) * T.qinv);
        }
#pragma unroll
        for (int q = DP_D; q < DP_RS; ++q) a.xq_out[ii * DP_RS + q] = a.xq[ii * DP_RS + q];
#endif
      }
    }
    __syncthreads();
  }
  if ((KIND == 2 || KIND == 3) && cur_rep >= 0) {
    double m = dmax;
#pragma unroll
    for (int off = 32; off >= 1; off >>= 1) m = fmax(m, __shfl_xor(m, off));
    if ((tid & 63) == 0) atomicMax(&a.dispmax[cur_rep], (unsigned long long)__double_as_longlong(m));
  }
}

extern "C" __global__ void __launch_bounds__(DP_TILE_THREADS) dp_tile_force(DpPairArgs a) { dp_tile_run<0>(a); }
extern "C" __global__ void __launch_bounds__(DP_TILE_THREADS) dp_tile_obs(DpPairArgs a) { dp_tile_run<1>(a); }
extern "C" __global__ void __launch_bounds__(DP_TILE_THREADS) dp_tile_gjf(DpPairArgs a) { dp_tile_run<2>(a); }
extern "C" __global__ void __launch_bounds__(DP_TILE_THREADS) dp_tile_lf(DpPairArgs a) { dp_tile_run<3>(a); }

// ---------------------------------------------------------------------------
// Integrator only, forces supplied in a.F: one thread per particle, no
// neighbour list.  Used for the free functions (integrator.py:5-40, no wrap,
// a.wrap = 0) and as the second half of the split step (force kernel, then
// this kernel with the wrap, the skin-displacement maximum and the mixed
// records).
// ---------------------------------------------------------------------------
template <int INTEG>
__device__ __forceinline__ void dp_integrate_body(const DpPairArgs& a) {
  const i64 i = (i64)blockIdx.x * blockDim.x + threadIdx.x;
  const bool valid = i < a.N;
  const i64 ii = valid ? i : 0;
  const int r = a.rep[ii];
  const DpRep& P = a.reps[r];
  double sdisp = 0.0;
  if (valid) {
#pragma clang fp contract(off)
    const double m = a.mass[ii];
    const double dt = P.dt;
    const u64 step = a.ctl->step_index;
    const u32 pid = (u32)a.orig[ii];
    double xn[DP_D];
    if (INTEG == 1) {
      const double alpha = m / P.tau;
      const double bvar = 2.0 * alpha * P.T_ * dt;
      const double sq = sqrt(bvar);
      const double qq = alpha * dt / 2.0 / m;
      const double num = 1.0 - qq;
      const double den = 1.0 + qq;
      const double ca = num / den;
      const double cb = 1.0 / den;
#pragma unroll
      for (int d = 0; d < DP_D; ++d) {
        const double z = a.xi ? a.xi[(i64)pid * DP_D + d] : dp_normal(a.seed, step, pid, d);
        const double nb = sq * z;
        const double v0 = a.v[ii * DP_D + d];
        const double b0 = a.beta[ii * DP_D + d];
        const double v1 = ca * v0 + cb * dt * a.F[ii * DP_D + d] / m + cb * (b0 + nb) / 2.0 / m;
        xn[d] = a.x[ii * DP_D + d] + v1 * dt;
        a.v[ii * DP_D + d] = v1;
        a.beta[ii * DP_D + d] = nb;
      }
    } else {
#pragma unroll
      for (int d = 0; d < DP_D; ++d) {
        const double v1 = a.v[ii * DP_D + d] + dt * a.F[ii * DP_D + d] / m;
        xn[d] = a.x[ii * DP_D + d] + dt * v1;
        a.v[ii * DP_D + d] = v1;
      }
    }
#pragma unroll
    for (int d = 0; d < DP_D; ++d) {
      if (a.wrap) {
        const double sh = floor(xn[d] / P.L[d]);
        a.img[ii * DP_D + d] += (i32)sh;
        xn[d] = xn[d] - sh * P.L[d];
      }
      a.x_out[ii * DP_D + d] = xn[d];
      if (a.wrap) {
        double tq = dp_sub(xn[d], a.xref[ii * DP_D + d]);
        tq = dp_minimage_gt(tq, P.L[d], P.halfL[d]);
        sdisp = (d == 0) ? dp_mul(tq, tq) : dp_add(sdisp, dp_mul(tq, tq));
      }
    }
    if (sdisp != sdisp) sdisp = __longlong_as_double(0x7ff0000000000000ll);
#if DP_MODE == 1
    if (a.wrap) {
      const u32 pack = (u32)a.cellpack[ii];
#pragma unroll
      for (int d = 0; d < DP_D; ++d) {
        const u32 c = (pack >> (DP_PACK_BITS * d)) & ((1u << DP_PACK_BITS) - 1u);
        const double cen = ((double)c + 0.5) * P.w[d];
        a.xq_out[ii * DP_RS + d] = (i32)rint(dp_minimage_gt(xn[d] - cen, P.L[d], P.halfL[d]) * a.tab->qinv);
      }
#pragma unroll
      for (int q = DP_D; q < DP_RS; ++q) a.xq_out[ii * DP_RS + q] = a.xq[ii * DP_RS + q];
    }
#endif
  }
  if (!a.wrap) return;
  // skin-displacement maximum: one atomic per wavefront when the wavefront
  // lies in one replica, per lane otherwise
  const int lane = threadIdx.x & 63;
  const unsigned long long b = __ballot(valid);
  if (b == 0ull) return;
  const int first = __ffsll((unsigned long long)b) - 1;
  const int rf = __shfl(r, first);
  if (__all((!valid) || r == rf)) {
    double mx = valid ? sdisp : 0.0;
#pragma unroll
    for (int off = 32; off >= 1; off >>= 1) mx = fmax(mx, __shfl_xor(mx, off));
    if (lane == first) atomicMax(&a.dispmax[rf], (unsigned long long)__double_as_longlong(mx));
  } else if (valid) {
    atomicMax(&a.dispmax[r], (unsigned long long)__double_as_longlong(sdisp));
  }
}

extern "C" __global__ void __launch_bounds__(256) dp_integrate_gjf(DpPairArgs a) { dp_integrate_body<1>(a); }
extern "C" __global__ void __launch_bounds__(256) dp_integrate_lf(DpPairArgs a) { dp_integrate_body<2>(a); }
