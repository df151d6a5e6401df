// dp_kernels.h -- hipcc-built pipeline kernels (templated on the spatial
// dimension D): integer cell binning (atomic counts, prefix scan, scatter),
// deterministic in-cell ordering, particle reordering, neighbour-list emit,
// skin check, per-step control, and layout conversions at the C-ABI.
// Included once, by dp_lib.hip.
#pragma once
#include "dp_device.h"

#define DP_BLOCK 256

// Pack bits per dimension for the cell coordinates kept with each particle.
template <int D> struct DpPack {
  static constexpr int BITS = (D == 1) ? 31 : 32 / D;
  static constexpr u32 MASK = (BITS >= 31) ? 0x7fffffffu : ((1u << BITS) - 1u);
};

// Grid-stride helpers: pipeline kernels are launched with a bounded grid so
// that a gated (skipped) launch costs one short wave of workgroups.
#define DP_FOR(i, n) for (i64 i = (i64)blockIdx.x * blockDim.x + threadIdx.x; i < (n); i += (i64)gridDim.x * blockDim.x)

__global__ void k_fill_i32(i32* p, i64 n, i32 v, const DpCtl* gate) {
  if (gate && !gate->rebuild) return;
  DP_FOR(i, n) p[i] = v;
}

__global__ void k_fill_u64(u64* p, i64 n, u64 v) {
  DP_FOR(i, n) p[i] = v;
}

// ---------------------------------------------------------------------------
// Control: decide whether this step rebuilds (neighbor_list.py:183-201 on the
// maximum computed by the previous integrator pass), reset the maximum, and
// take the RNG counter for this step.
// ---------------------------------------------------------------------------
__global__ void k_prologue(DpCtl* ctl, u64* dispmax, const DpRep* reps, int R, i64* nl_updates,
                           int advance_step) {
  if (blockIdx.x != 0 || threadIdx.x != 0) return;
  int old = 0;
  for (int r = 0; r < R; ++r) {
    const double m = __longlong_as_double((long long)dispmax[r]);
    if (m > reps[r].skin_half_sq) old = 1;
  }
  const int need = old | ctl->force_rebuild;
  ctl->rebuild = need;
  ctl->force_rebuild = 0;
  if (need) ctl->max_count = 0;
  // only skin-triggered rebuilds are counted here; requested ones are counted
  // by the caller that requested them
  if (old) {
    for (int r = 0; r < R; ++r) nl_updates[r] += 1;
  }
  for (int r = 0; r < R; ++r) dispmax[r] = 0ull;
  if (advance_step) {
    ctl->step_index = ctl->step_counter;
    ctl->step_counter += 1;
  }
}

__global__ void k_advance_step(DpCtl* ctl) {
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    ctl->step_index = ctl->step_counter;
    ctl->step_counter += 1;
  }
}

__global__ void k_request_rebuild(DpCtl* ctl) {
  if (blockIdx.x == 0 && threadIdx.x == 0) ctl->force_rebuild = 1;
}

// Skin check from the current positions (used outside the fused step, e.g.
// after the host uploaded new positions).
template <int D>
__global__ void k_skin_check(const double* x, const double* xref, const i32* rep, const DpRep* reps, i64 N,
                             u64* dispmax) {
  DP_FOR(i, N) {
    const DpRep& R = reps[rep[i]];
    double s = 0.0;
#pragma unroll
    for (int d = 0; d < D; ++d) {
      double t = dp_sub(x[i * D + d], xref[i * D + d]);
      t = dp_minimage_gt(t, R.L[d], R.halfL[d]);
      s = (d == 0) ? dp_mul(t, t) : dp_add(s, dp_mul(t, t));
    }
    if (s != s) s = __longlong_as_double(0x7ff0000000000000ll);  // NaN -> +inf: forces a rebuild
    atomicMax(&dispmax[rep[i]], (unsigned long long)__double_as_longlong(s));
  }
}

// ---------------------------------------------------------------------------
// Binning: cell of each particle from its wrapped FP64 position; atomic count
// gives a provisional rank inside the cell.
// ---------------------------------------------------------------------------
template <int D>
__global__ void k_bin(const double* x, const i32* rep, const DpRep* reps, i64 N, i32* cell_of, i32* rank,
                      i32* count, DpCtl* ctl) {
  if (!ctl->rebuild) return;
  DP_FOR(i, N) {
    const DpRep& R = reps[rep[i]];
    i64 lin = 0, stride = 1;
    bool finite = true;
#pragma unroll
    for (int d = 0; d < D; ++d) {
      double xd = x[i * D + d];
      if (!(xd - xd == 0.0)) { finite = false; xd = 0.0; }
      xd = xd - floor(xd / R.L[d]) * R.L[d];
      int c = (int)floor(xd * R.inv_w[d]);
      if (c < 0) c = 0;
      if (c >= R.nc[d]) c = R.nc[d] - 1;
      lin += (i64)c * stride;
      stride *= R.nc[d];
    }
    if (!finite) ctl->nonfinite = 1;
    const i32 cid = (i32)(R.c0 + lin);
    cell_of[i] = cid;
    rank[i] = atomicAdd(&count[cid], 1);
  }
}

// Exclusive scan of cell counts: cstart[c] = sum_{c' < c} count[c'];
// cstart[C] = N.  Three passes over blocks of DP_SCAN_TILE entries.
#define DP_SCAN_TILE 1024

__device__ __forceinline__ i32 dp_block_excl_scan(i32 v, i32* sh, i32& total) {
  // 256-thread inclusive scan in LDS (Hillis-Steele), returns exclusive
  const int t = threadIdx.x;
  sh[t] = v;
  __syncthreads();
  for (int off = 1; off < DP_BLOCK; off <<= 1) {
    i32 add = (t >= off) ? sh[t - off] : 0;
    __syncthreads();
    sh[t] += add;
    __syncthreads();
  }
  total = sh[DP_BLOCK - 1];
  const i32 incl = sh[t];
  __syncthreads();
  return incl - v;
}

__global__ void k_scan_reduce(const i32* count, i64 C, i32* tile_sums, const DpCtl* ctl) {
  if (!ctl->rebuild) return;
  __shared__ i32 sh[DP_BLOCK];
  for (i64 tile = blockIdx.x; tile * DP_SCAN_TILE < C; tile += gridDim.x) {
    const i64 base = tile * DP_SCAN_TILE + (i64)threadIdx.x * 4;
    i32 s = 0;
#pragma unroll
    for (int k = 0; k < 4; ++k)
      if (base + k < C) s += count[base + k];
    i32 total;
    dp_block_excl_scan(s, sh, total);
    if (threadIdx.x == 0) tile_sums[tile] = total;
  }
}

__global__ void k_scan_top(i32* tile_sums, i64 ntiles, const DpCtl* ctl) {
  if (!ctl->rebuild) return;
  __shared__ i32 sh[DP_BLOCK];
  i32 carry = 0;
  for (i64 base = 0; base < ntiles; base += DP_BLOCK) {
    const i64 k = base + threadIdx.x;
    const i32 v = (k < ntiles) ? tile_sums[k] : 0;
    i32 total;
    const i32 ex = dp_block_excl_scan(v, sh, total);
    if (k < ntiles) tile_sums[k] = carry + ex;
    carry += total;
    __syncthreads();
  }
}

__global__ void k_scan_down(const i32* count, i64 C, const i32* tile_offsets, i32* cstart, i64 N,
                            const DpCtl* ctl) {
  if (!ctl->rebuild) return;
  __shared__ i32 sh[DP_BLOCK];
  for (i64 tile = blockIdx.x; tile * DP_SCAN_TILE < C; tile += gridDim.x) {
    const i64 base = tile * DP_SCAN_TILE + (i64)threadIdx.x * 4;
    i32 v[4];
    i32 s = 0;
#pragma unroll
    for (int k = 0; k < 4; ++k) {
      v[k] = (base + k < C) ? count[base + k] : 0;
      s += v[k];
    }
    i32 total;
    i32 ex = dp_block_excl_scan(s, sh, total) + tile_offsets[tile];
#pragma unroll
    for (int k = 0; k < 4; ++k) {
      if (base + k < C) cstart[base + k] = ex;
      ex += v[k];
    }
  }
  if (blockIdx.x == 0 && threadIdx.x == 0) cstart[C] = (i32)N;
}

__global__ void k_scatter(const i32* cell_of, const i32* rank, const i32* cstart, i64 N, i32* tmp,
                          const DpCtl* ctl) {
  if (!ctl->rebuild) return;
  DP_FOR(i, N) tmp[cstart[cell_of[i]] + rank[i]] = (i32)i;
}

// In-cell order by original particle id (deterministic regardless of the
// atomic ranks): the new position of a member is the number of members of
// its cell with a smaller id.
__global__ void k_cellsort(const i32* tmp, const i32* cell_of, const i32* cstart, const i32* orig, i64 N,
                           i32* perm, const DpCtl* ctl) {
  if (!ctl->rebuild) return;
  DP_FOR(q, N) {
    const i32 p = tmp[q];
    const i32 c = cell_of[p];
    const i32 lo = cstart[c], hi = cstart[c + 1];
    const i32 key = orig[p];
    i32 rnk = 0;
    for (i32 k = lo; k < hi; ++k) rnk += (orig[tmp[k]] < key) ? 1 : 0;
    perm[lo + rnk] = p;
  }
}

// Gather every per-particle array into the new order (into scratch), then
// copy back.  Scratch layout: x | v | beta (N*D doubles each), mass (N),
// img (N*D i32), type | orig | cell (N i32 each).
template <int D>
__global__ void k_permute(const i32* perm, i64 N, const double* x, const double* v, const double* beta,
                          const double* mass, const i32* img, const i32* type, const i32* orig,
                          const i32* cell_of, char* scratch, const DpCtl* ctl) {
  if (!ctl->rebuild) return;
  double* sx = (double*)scratch;
  double* sv = sx + N * D;
  double* sb = sv + N * D;
  double* sm = sb + N * D;
  i32* si = (i32*)(sm + N);
  i32* st = si + N * D;
  i32* so = st + N;
  i32* sc = so + N;
  DP_FOR(q, N) {
    const i64 p = perm[q];
#pragma unroll
    for (int d = 0; d < D; ++d) {
      sx[q * D + d] = x[p * D + d];
      sv[q * D + d] = v[p * D + d];
      sb[q * D + d] = beta[p * D + d];
      si[q * D + d] = img[p * D + d];
    }
    sm[q] = mass[p];
    st[q] = type[p];
    so[q] = orig[p];
    sc[q] = cell_of[p];
  }
}

// Copy back from scratch, record the reference positions (Simulation.py:381
// positions_neighbour_list = positions.copy()), pack cell coordinates, and
// write the mixed-precision records (FP32 position relative to the cell
// centre; type bits in the slot after the coordinates).
template <int D>
__global__ void k_unpermute_store(i64 N, double* x, double* v, double* beta, double* mass, i32* img, i32* type,
                                  i32* orig, i32* cell_of, const char* scratch, double* xref, i32* cellpack,
                                  float* xs, int rs, const i32* rep, const DpRep* reps, const DpCtl* ctl) {
  if (!ctl->rebuild) return;
  const double* sx = (const double*)scratch;
  const double* sv = sx + N * D;
  const double* sb = sv + N * D;
  const double* sm = sb + N * D;
  const i32* si = (const i32*)(sm + N);
  const i32* st = si + N * D;
  const i32* so = st + N;
  const i32* sc = so + N;
  DP_FOR(q, N) {
    const DpRep& R = reps[rep[q]];
    const i32 cid = sc[q];
    i64 lin = (i64)cid - R.c0;
    u32 pack = 0;
#pragma unroll
    for (int d = 0; d < D; ++d) {
      const i64 c = lin % R.nc[d];
      lin /= R.nc[d];
      pack |= ((u32)c) << (DpPack<D>::BITS * d);
      const double xd = sx[q * D + d];
      x[q * D + d] = xd;
      xref[q * D + d] = xd;
      v[q * D + d] = sv[q * D + d];
      beta[q * D + d] = sb[q * D + d];
      img[q * D + d] = si[q * D + d];
      if (xs) {
        const double cen = ((double)c + 0.5) * R.w[d];
        xs[q * rs + d] = (float)dp_minimage_gt(xd - cen, R.L[d], R.halfL[d]);
      }
    }
    mass[q] = sm[q];
    type[q] = st[q];
    orig[q] = so[q];
    cell_of[q] = cid;
    cellpack[q] = (i32)pack;
    if (xs) {
      if (D < rs) xs[q * rs + D] = __int_as_float(st[q]);
      for (int k = D + 1; k < rs; ++k) xs[q * rs + k] = 0.f;
    }
  }
}

// Mixed records from the current positions (after host uploads).
template <int D>
__global__ void k_make_records(i64 N, const double* x, const i32* type, const i32* cellpack, float* xs, int rs,
                               const i32* rep, const DpRep* reps) {
  DP_FOR(q, N) {
    const DpRep& R = reps[rep[q]];
    const u32 pack = (u32)cellpack[q];
#pragma unroll
    for (int d = 0; d < D; ++d) {
      const u32 c = (pack >> (DpPack<D>::BITS * d)) & DpPack<D>::MASK;
      const double cen = ((double)c + 0.5) * R.w[d];
      xs[q * rs + d] = (float)dp_minimage_gt(x[q * D + d] - cen, R.L[d], R.halfL[d]);
    }
    if (D < rs) xs[q * rs + D] = __int_as_float(type[q]);
    for (int k = D + 1; k < rs; ++k) xs[q * rs + k] = 0.f;
  }
}

// ---------------------------------------------------------------------------
// Neighbour-list emit: for each particle, the stencil cells in the
// reference's order (neighbor_list.py:150-157, x fastest, for order == 0;
// neighbor_list.py:82-86, last dimension fastest, for order == 1), members in
// ascending original id, kept iff the FP64 distance test of
// neighbor_list.py:24-33 holds (evaluated as s < threshold, which is exactly
// sqrt(s) < cutoff).  Entries are slot << 2D | stencil code.
// ---------------------------------------------------------------------------
template <int D>
__global__ void __launch_bounds__(DP_BLOCK) k_nl_emit(const double* x, const i32* cellpack, const i32* cstart,
                                                      const i32* rep, const DpRep* reps, i64 N, i64 Ns, int K,
                                                      int order, u32* nl, i32* nlcnt, DpCtl* ctl) {
  if (!ctl->rebuild) return;
  DP_FOR(i, N) {
    const DpRep& R = reps[rep[i]];
    double xi[D];
    int ci[D];
    const u32 pack = (u32)cellpack[i];
#pragma unroll
    for (int d = 0; d < D; ++d) {
      xi[d] = x[i * D + d];
      ci[d] = (int)((pack >> (DpPack<D>::BITS * d)) & DpPack<D>::MASK);
    }
    int n3 = 1;
#pragma unroll
    for (int d = 0; d < D; ++d) n3 *= 3;
    int cnt = 0;
    for (int st = 0; st < n3; ++st) {
      int s[D];
      int rem = st;
      bool skip = false;
      if (order == 0) {
#pragma unroll
        for (int d = 0; d < D; ++d) { s[d] = rem % 3 - 1; rem /= 3; }
      } else {
#pragma unroll
        for (int d = D - 1; d >= 0; --d) { s[d] = rem % 3 - 1; rem /= 3; }
      }
      i64 lin = 0, stride = 1;
      u32 code = 0;
#pragma unroll
      for (int d = 0; d < D; ++d) {
        if (R.nc[d] == 1 && s[d] != 0) skip = true;
        int c = ci[d] + s[d];
        if (c < 0) c += R.nc[d];
        else if (c >= R.nc[d]) c -= R.nc[d];
        lin += (i64)c * stride;
        stride *= R.nc[d];
        code |= (u32)(s[d] + 1) << (2 * d);
      }
      if (skip) continue;
      const i64 cell = R.c0 + lin;
      const i32 lo = cstart[cell], hi = cstart[cell + 1];
      for (i32 j = lo; j < hi; ++j) {
        if ((i64)j == i) continue;
        double ss = 0.0;
#pragma unroll
        for (int d = 0; d < D; ++d) {
          double t = dp_sub(xi[d], x[(i64)j * D + d]);
          t = dp_minimage_lt(t, R.L[d], R.halfL[d]);
          ss = (d == 0) ? dp_mul(t, t) : dp_add(ss, dp_mul(t, t));
        }
        if (ss < R.nl_thresh) {
          if (cnt < K) nl[(i64)cnt * Ns + i] = ((u32)j << (2 * D)) | code;
          ++cnt;
        }
      }
    }
    nlcnt[i] = cnt;
    if (cnt > K) ctl->overflow = 1;
    atomicMax(&ctl->max_count, cnt);
  }
}

// ---------------------------------------------------------------------------
// Layout conversion at the C-ABI: slot order <-> original particle order.
// ---------------------------------------------------------------------------
template <typename T>
__global__ void k_to_orig(const T* src, T* dst, const i32* orig, i64 N, int W) {
  DP_FOR(q, N) {
    const i64 o = orig[q];
    for (int k = 0; k < W; ++k) dst[o * W + k] = src[q * W + k];
  }
}

template <typename T>
__global__ void k_from_orig(const T* src, T* dst, const i32* orig, i64 N, int W) {
  DP_FOR(q, N) {
    const i64 o = orig[q];
    for (int k = 0; k < W; ++k) dst[q * W + k] = src[o * W + k];
  }
}

// Neighbour list in the reference layout: row orig[i], entries orig[j],
// -1 padded to K (neighbor_list.py:14-17).
__global__ void k_nl_export(const u32* nl, const i32* nlcnt, const i32* orig, i64 N, i64 Ns, int K, int sb,
                            i32* out) {
  DP_FOR(i, N) {
    const i64 row = orig[i];
    const int cnt = nlcnt[i] < K ? nlcnt[i] : K;
    for (int k = 0; k < K; ++k)
      out[row * K + k] = (k < cnt) ? orig[nl[(i64)k * Ns + i] >> sb] : -1;
  }
}

// Deterministic per-replica sums (fixed assignment of elements to threads and
// a fixed-shape tree), one workgroup per replica.
__global__ void k_reduce_rep(const double* vals, const DpRep* reps, double* out) {
  __shared__ double sh[DP_BLOCK];
  const DpRep& R = reps[blockIdx.x];
  double s = 0.0;
  for (i64 k = threadIdx.x; k < R.n; k += DP_BLOCK) s += vals[R.p0 + k];
  sh[threadIdx.x] = s;
  __syncthreads();
  for (int off = DP_BLOCK / 2; off > 0; off >>= 1) {
    if ((int)threadIdx.x < off) sh[threadIdx.x] += sh[threadIdx.x + off];
    __syncthreads();
  }
  if (threadIdx.x == 0) out[blockIdx.x] = sh[0];
}
