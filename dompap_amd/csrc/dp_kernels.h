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

// The rebuild pipeline's particles: with plist (batches), the ctl->pblk_n
// blocks [start, end) of the replicas that rebuild in this step (DP_BLOCK
// particles each, built by k_rebuild_list); else every particle (the kernels
// then skip those of replicas that keep their list).  f(i, valid) is called by
// every lane of a workgroup together (valid lanes form a prefix).
template <class F>
__device__ __forceinline__ void dp_for_rebuild(const int2* plist, const DpCtl* ctl, i64 N, F&& f) {
  if (plist) {
    const int nb = ctl->pblk_n;
    for (int b = blockIdx.x; b < nb; b += gridDim.x) {
      const int2 be = plist[b];
      const i64 i = (i64)be.x + threadIdx.x;
      f(i, i < (i64)be.y);
    }
  } else {
    for (i64 base = (i64)blockIdx.x * blockDim.x; base < N; base += (i64)gridDim.x * blockDim.x) {
      const i64 i = base + threadIdx.x;
      f(i, i < N);
    }
  }
}

// ---------------------------------------------------------------------------
// Control: decide which replicas rebuild this step (neighbor_list.py:183-201:
// the previous integrator pass raised need[r] iff some particle of replica r
// moved by more than half the skin since r's last build, or to a non-finite
// position), clear the flags, and take the RNG counter for this step.
// rflag[0..R) = need, rflag[R..2R) = rebuild of the current step; only the
// replicas whose trigger fired rebuild (all of them on a host request).
// ---------------------------------------------------------------------------
// Called by every thread of one workgroup; returns whether any replica
// rebuilds.  Only skin-triggered rebuilds are counted here; requested ones
// are counted by the caller that requested them.
__device__ __forceinline__ int dp_decide_rebuilds(i32* rflag, int R, i64* nl_updates, int forced) {
  __shared__ int any_s;
  if (threadIdx.x == 0) any_s = 0;
  __syncthreads();
  int any = 0;
  for (int r = threadIdx.x; r < R; r += blockDim.x) {
    const int nd = rflag[r];
    const int rb = (nd | forced) ? 1 : 0;
    rflag[R + r] = rb;
    rflag[r] = 0;
    if (nd) nl_updates[r] += 1;
    any |= rb;
  }
  if (any) any_s = 1;
  __syncthreads();
  return any_s;
}

__device__ __forceinline__ void dp_clear_stats(DpCtl* ctl) {
  for (int k = threadIdx.x; k < DP_STAT_SLOTS * DP_STAT_STRIDE; k += blockDim.x) ctl->slot[k] = 0;
}

__global__ void k_prologue(DpCtl* ctl, i32* rflag, int R, i64* nl_updates, int advance_step) {
  if (blockIdx.x != 0) return;
  const int forced = ctl->force_rebuild;
  const int any = dp_decide_rebuilds(rflag, R, nl_updates, forced);
  if (any) dp_clear_stats(ctl);
  if (threadIdx.x != 0) return;
  ctl->rebuild = any;
  ctl->force_rebuild = 0;
  ctl->need_rebuild = 0;
  if (advance_step) {
    ctl->step_index = ctl->step_counter;
    ctl->step_counter += 1;
  }
}

// Step blocks (a captured graph of `nb` steps).  k_block_begin clears the
// block's per-step flags and fixes the block's first step index and the last
// one it may run (blk_left steps from here, set by the host).  perstep = 0:
// the rebuild pipeline follows once, at the block's start, and k_block_begin
// takes the place of the first step's prologue.  perstep = 1: every step
// starts with k_step_begin and the (gated) pipeline.
__global__ void k_block_begin(DpCtl* ctl, i64* nl_updates, i32* rflag, int R, int perstep) {
  if (blockIdx.x != 0) return;
  const int tid = threadIdx.x;
  const int forced = ctl->force_rebuild;
  for (int k = tid; k < DP_BLK_MAX; k += blockDim.x) ctl->blk_need[k] = 0;
  int any = 0;
  if (!perstep) {
    any = dp_decide_rebuilds(rflag, R, nl_updates, forced);
    if (any) dp_clear_stats(ctl);
  }
  if (tid != 0) return;
  ctl->blk_perstep = perstep;
  if (!perstep) {
    ctl->rebuild = any;
    ctl->force_rebuild = 0;
    ctl->need_rebuild = 0;
  }
  ctl->blk_base = ctl->step_counter;
  ctl->blk_target = ctl->step_counter + ctl->blk_left;
  ctl->blk_done = 0;
}

// Step `pos` of a per-step block: the rebuild decision of that step (none
// for a step beyond the block's target, whose kernels return at once).
__global__ void k_step_begin(DpCtl* ctl, i64* nl_updates, i32* rflag, int R, int pos) {
  if (blockIdx.x != 0) return;
  if (ctl->blk_base + (u64)pos >= ctl->blk_target) {
    if (threadIdx.x == 0) ctl->rebuild = 0;
    return;
  }
  const int forced = ctl->force_rebuild;
  const int any = dp_decide_rebuilds(rflag, R, nl_updates, forced);
  if (any) dp_clear_stats(ctl);
  if (threadIdx.x != 0) return;
  ctl->rebuild = any;
  ctl->force_rebuild = 0;
  ctl->need_rebuild = 0;
}

// End of a step block: the steps that ran (below the target; with perstep =
// 0 also up to and including the first one that raised its flag), the new
// step counter and the rebuild request for the next step.
__global__ void k_block_end(DpCtl* ctl, int nb) {
  if (blockIdx.x != 0 || threadIdx.x != 0) return;
  int done = 0, need = 0;
  const int perstep = ctl->blk_perstep;
  for (int p = 0; p < nb; ++p) {
    if (ctl->blk_base + (u64)p >= ctl->blk_target) break;
    ++done;
    if (!perstep && ctl->blk_need[p]) {
      need = 1;
      break;
    }
  }
  ctl->blk_done = done;
  ctl->step_counter = ctl->blk_base + (u64)done;
  if (done > 0) ctl->step_index = ctl->step_counter - 1;
  if (need) ctl->need_rebuild = 1;
  ctl->blk_perstep = 0;
}

__global__ void k_block_left(DpCtl* ctl, i64 left) {
  if (blockIdx.x == 0 && threadIdx.x == 0) ctl->blk_left = (u64)left;
}

__global__ void k_advance_step(DpCtl* ctl) {
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    ctl->step_index = ctl->step_counter;
    ctl->step_counter += 1;
  }
}

__global__ void k_clear_need(DpCtl* ctl, i32* rflag, int R) {
  if (blockIdx.x != 0) return;
  for (int r = threadIdx.x; r < R; r += blockDim.x) rflag[r] = 0;
  if (threadIdx.x == 0) ctl->need_rebuild = 0;
}

__global__ void k_request_rebuild(DpCtl* ctl) {
  if (blockIdx.x == 0 && threadIdx.x == 0) ctl->force_rebuild = 1;
}

// Skin check from the current positions (used outside the fused step, e.g.
// after the host uploaded new positions).
template <int D>
__global__ void k_skin_check(const double* x, const double* xref, const i32* rep, const DpRep* reps, i64 N,
                             DpCtl* ctl, i32* need) {
  bool trip = false;
  DP_FOR(i, N) {
    const int r = rep[i];
    const DpRep& R = reps[r];
    double s = 0.0;
#pragma unroll
    for (int d = 0; d < D; ++d) {
      double t = dp_sub(x[i * D + d], xref[i * D + d]);
      t = dp_minimage_gt(t, R.L[d], R.halfL[d]);
      s = (d == 0) ? dp_mul(t, t) : dp_add(s, dp_mul(t, t));
    }
    const bool ti = dp_skin_trips(s, R.skin_half_sq);
    dp_raise_replica(ti, need, r);
    trip |= ti;
  }
  dp_raise_rebuild(trip, &ctl->need_rebuild);
}

// A particle's D doubles, two at a time (8-byte aligned pairs: 16-byte
// loads and stores, fewer memory instructions per cache line).
typedef double dp_d2a8 __attribute__((ext_vector_type(2), aligned(8)));
template <int D>
__device__ __forceinline__ void dp_ldD(const double* p, double (&v)[D]) {
#pragma unroll
  for (int d = 0; d + 1 < D; d += 2) {
    const dp_d2a8 t = *(const dp_d2a8*)(p + d);
    v[d] = t.x;
    v[d + 1] = t.y;
  }
  if (D & 1) v[D - 1] = p[D - 1];
}
template <int D>
__device__ __forceinline__ void dp_stD(double* p, const double (&v)[D]) {
#pragma unroll
  for (int d = 0; d + 1 < D; d += 2) {
    dp_d2a8 t;
    t.x = v[d];
    t.y = v[d + 1];
    *(dp_d2a8*)(p + d) = t;
  }
  if (D & 1) p[D - 1] = v[D - 1];
}

// ---------------------------------------------------------------------------
// Binning: cell of each particle from its wrapped FP64 position; atomic count
// gives a provisional rank inside the cell.
// ---------------------------------------------------------------------------
// A replica that does not rebuild this step (rb[r] == 0) keeps its cells:
// its particles are counted in the cell they were sorted into at its last
// build (cell_of), so the scan reproduces its cell starts, and the later
// pipeline kernels leave its particles, records and list untouched.
template <int D>
__global__ void k_bin(const double* x, const i32* rep, const DpRep* reps, i64 N, i32* cell_of, i32* rank,
                      i32* count, DpCtl* ctl, const i32* rb, const int2* plist) {
  if (!ctl->rebuild) return;
  if (blockIdx.x == 0 && threadIdx.x == 0) ctl->far_tiles = 0;  // set again by this rebuild's emit
  const int lane = threadIdx.x & 63;
  dp_for_rebuild(plist, ctl, N, [&](const i64 i, const bool valid) {  // valid lanes form a prefix of the wavefront
    i32 cid = -1;
    if (valid && !plist && !rb[rep[i]]) {
      cid = cell_of[i];
    } else if (valid) {
      const DpRep& R = reps[rep[i]];
      int c[D];
      bool finite = true;
      double xi[D];
      dp_ldD<D>(x + i * D, xi);
#pragma unroll
      for (int d = 0; d < D; ++d) {
        double xd = xi[d];
        if (!(xd - xd == 0.0)) { finite = false; xd = 0.0; }
        xd = xd - floor(xd / R.L[d]) * R.L[d];
        int cd = (int)floor(xd * R.inv_w[d]);
        if (cd < 0) cd = 0;
        if (cd >= R.nc[d]) cd = R.nc[d] - 1;
        c[d] = cd;
      }
      if (!finite) ctl->nonfinite = 1;
      cid = (i32)(R.c0 + dp_linear<D>(c, R.nc, R.order));
      cell_of[i] = cid;
    }
    // Particles are in cell order from the previous build, so neighbouring
    // lanes mostly share a cell: one atomic per run of equal cells, the
    // run's lanes take consecutive ranks after it.
    const i32 prev = __shfl_up(cid, 1);
    const bool head = valid && (lane == 0 || prev != cid);
    const u64 heads = __ballot(head);
    const u64 vmask = __ballot(valid);
    const u64 le = (lane == 63) ? ~0ull : ((1ull << (lane + 1)) - 1ull);
    const u64 hle = heads & le;
    const int start = hle ? 63 - __clzll(hle) : 0;
    int b = 0;
    if (head) {
      const u64 above = heads & ~le;
      const int next = above ? __ffsll((unsigned long long)above) - 1 : 64 - __clzll(vmask);
      b = atomicAdd(&count[cid], next - lane);
    }
    b = __shfl(b, start);
    if (valid) rank[i] = b + (lane - start);
  });
}

// Single-workgroup exclusive scan of the cell counts: blocks of 16384 cells,
// 16 per thread (four 16-byte loads), wavefront scans combined through LDS, a
// running carry.  cstart[c] = sum of the counts before c, cstart[C] = N.  The
// counts are cleared for the next build as they are read.
#define DP_SCAN_THREADS 1024
#define DP_SCAN_PER 16
__global__ void __launch_bounds__(DP_SCAN_THREADS) k_scan_single(i32* count, i64 C, i32* cstart, i64 N,
                                                                  const DpCtl* ctl) {
  if (!ctl->rebuild) return;
  __shared__ i32 wsum[DP_SCAN_THREADS / 64];
  __shared__ i32 btotal;
  const int t = threadIdx.x;
  const int lane = t & 63;
  const int w = t >> 6;
  constexpr int NW = DP_SCAN_THREADS / 64;
  i32 carry = 0;
  for (i64 b0 = 0; b0 < C; b0 += DP_SCAN_PER * DP_SCAN_THREADS) {
    const i64 c0 = b0 + DP_SCAN_PER * (i64)t;
    i32 v[DP_SCAN_PER];
    if (c0 + DP_SCAN_PER <= C) {
#pragma unroll
      for (int k = 0; k < DP_SCAN_PER; k += 4) {
        const int4 q = *(const int4*)(count + c0 + k);
        v[k] = q.x; v[k + 1] = q.y; v[k + 2] = q.z; v[k + 3] = q.w;
      }
    } else {
#pragma unroll
      for (int k = 0; k < DP_SCAN_PER; ++k) v[k] = (c0 + k < C) ? count[c0 + k] : 0;
    }
    i32 sthr = 0;
#pragma unroll
    for (int k = 0; k < DP_SCAN_PER; ++k) sthr += v[k];
    i32 incl = sthr;
#pragma unroll
    for (int off = 1; off < 64; off <<= 1) {
      const i32 y = __shfl_up(incl, off);
      if (lane >= off) incl += y;
    }
    if (lane == 63) wsum[w] = incl;
    __syncthreads();
    if (w == 0) {
      const i32 xw = (lane < NW) ? wsum[lane] : 0;
      i32 xi = xw;
#pragma unroll
      for (int off = 1; off < NW; off <<= 1) {
        const i32 y = __shfl_up(xi, off);
        if (lane >= off) xi += y;
      }
      if (lane < NW) wsum[lane] = xi - xw;  // exclusive prefix of the wavefront totals
      if (lane == NW - 1) btotal = xi;
    }
    __syncthreads();
    i32 ex = carry + wsum[w] + (incl - sthr);
    if (c0 + DP_SCAN_PER <= C) {
#pragma unroll
      for (int k = 0; k < DP_SCAN_PER; k += 4) {
        int4 o;
        o.x = ex; ex += v[k];
        o.y = ex; ex += v[k + 1];
        o.z = ex; ex += v[k + 2];
        o.w = ex; ex += v[k + 3];
        *(int4*)(cstart + c0 + k) = o;
        *(int4*)(count + c0 + k) = make_int4(0, 0, 0, 0);
      }
    } else {
#pragma unroll
      for (int k = 0; k < DP_SCAN_PER; ++k) {
        if (c0 + k < C) {
          cstart[c0 + k] = ex;
          ex += v[k];
          count[c0 + k] = 0;
        }
      }
    }
    carry += btotal;
    __syncthreads();  // wsum / btotal are rewritten by the next block
  }
  if (t == 0) cstart[C] = (i32)N;
}

// The same scan over many workgroups, for many cells: each workgroup sums
// its block of 16384 counts (k_scan_blocks), one workgroup scans the block
// sums (k_scan_top), and each workgroup then writes its block's starts from
// that offset (k_scan_apply), clearing the counts.
template <bool APPLY>
__device__ __forceinline__ void dp_scan_block(i32* count, i64 C, i32* cstart, i32* part, i64 N) {
  __shared__ i32 wsum[DP_SCAN_THREADS / 64];
  __shared__ i32 btotal;
  const int t = threadIdx.x;
  const int lane = t & 63;
  const int w = t >> 6;
  constexpr int NW = DP_SCAN_THREADS / 64;
  const i64 b0 = (i64)blockIdx.x * (DP_SCAN_PER * DP_SCAN_THREADS);
  const i64 c0 = b0 + DP_SCAN_PER * (i64)t;
  i32 v[DP_SCAN_PER];
  if (c0 + DP_SCAN_PER <= C) {
#pragma unroll
    for (int k = 0; k < DP_SCAN_PER; k += 4) {
      const int4 q = *(const int4*)(count + c0 + k);
      v[k] = q.x; v[k + 1] = q.y; v[k + 2] = q.z; v[k + 3] = q.w;
    }
  } else {
#pragma unroll
    for (int k = 0; k < DP_SCAN_PER; ++k) v[k] = (c0 + k < C) ? count[c0 + k] : 0;
  }
  i32 sthr = 0;
#pragma unroll
  for (int k = 0; k < DP_SCAN_PER; ++k) sthr += v[k];
  i32 incl = sthr;
#pragma unroll
  for (int off = 1; off < 64; off <<= 1) {
    const i32 y = __shfl_up(incl, off);
    if (lane >= off) incl += y;
  }
  if (lane == 63) wsum[w] = incl;
  __syncthreads();
  if (w == 0) {
    const i32 xw = (lane < NW) ? wsum[lane] : 0;
    i32 xi = xw;
#pragma unroll
    for (int off = 1; off < NW; off <<= 1) {
      const i32 y = __shfl_up(xi, off);
      if (lane >= off) xi += y;
    }
    if (lane < NW) wsum[lane] = xi - xw;
    if (lane == NW - 1) btotal = xi;
  }
  __syncthreads();
  if (!APPLY) {
    if (t == 0) part[blockIdx.x] = btotal;
    return;
  }
  i32 ex = part[blockIdx.x] + wsum[w] + (incl - sthr);
  if (c0 + DP_SCAN_PER <= C) {
#pragma unroll
    for (int k = 0; k < DP_SCAN_PER; k += 4) {
      int4 o;
      o.x = ex; ex += v[k];
      o.y = ex; ex += v[k + 1];
      o.z = ex; ex += v[k + 2];
      o.w = ex; ex += v[k + 3];
      *(int4*)(cstart + c0 + k) = o;
      *(int4*)(count + c0 + k) = make_int4(0, 0, 0, 0);
    }
  } else {
#pragma unroll
    for (int k = 0; k < DP_SCAN_PER; ++k) {
      if (c0 + k < C) {
        cstart[c0 + k] = ex;
        ex += v[k];
        count[c0 + k] = 0;
      }
    }
  }
  if (blockIdx.x == 0 && t == 0) cstart[C] = (i32)N;
}
__global__ void __launch_bounds__(DP_SCAN_THREADS) k_scan_blocks(i32* count, i64 C, i32* part, const DpCtl* ctl) {
  if (!ctl->rebuild) return;
  dp_scan_block<false>(count, C, nullptr, part, 0);
}
// Exclusive scan of the nb block sums in place (one workgroup, a running carry).
__global__ void __launch_bounds__(DP_SCAN_THREADS) k_scan_top(i32* part, i64 nb, const DpCtl* ctl) {
  if (!ctl->rebuild) return;
  __shared__ i32 wsum[DP_SCAN_THREADS / 64];
  __shared__ i32 btotal;
  const int t = threadIdx.x;
  const int lane = t & 63;
  const int w = t >> 6;
  constexpr int NW = DP_SCAN_THREADS / 64;
  i32 carry = 0;
  for (i64 b0 = 0; b0 < nb; b0 += DP_SCAN_THREADS) {
    const i64 b = b0 + t;
    const i32 x = (b < nb) ? part[b] : 0;
    i32 incl = x;
#pragma unroll
    for (int off = 1; off < 64; off <<= 1) {
      const i32 y = __shfl_up(incl, off);
      if (lane >= off) incl += y;
    }
    if (lane == 63) wsum[w] = incl;
    __syncthreads();
    if (w == 0) {
      const i32 xw = (lane < NW) ? wsum[lane] : 0;
      i32 xi = xw;
#pragma unroll
      for (int off = 1; off < NW; off <<= 1) {
        const i32 y = __shfl_up(xi, off);
        if (lane >= off) xi += y;
      }
      if (lane < NW) wsum[lane] = xi - xw;
      if (lane == NW - 1) btotal = xi;
    }
    __syncthreads();
    if (b < nb) part[b] = carry + wsum[w] + (incl - x);
    carry += btotal;
    __syncthreads();
  }
}
__global__ void __launch_bounds__(DP_SCAN_THREADS) k_scan_apply(i32* count, i64 C, i32* cstart, const i32* part,
                                                                 i64 N, const DpCtl* ctl) {
  if (!ctl->rebuild) return;
  dp_scan_block<true>(count, C, cstart, const_cast<i32*>(part), N);
}

// Particle i to position cstart[cell] + rank in the cell-grouped order, with
// its original id alongside (keys[pos]) for the in-cell sort.
__global__ void k_scatter(const i32* cell_of, const i32* rank, const i32* cstart, const i32* orig, i64 N, i32* tmp,
                          i32* keys, const DpCtl* ctl, const i32* rep, const i32* rb, const int2* plist) {
  if (!ctl->rebuild) return;
  dp_for_rebuild(plist, ctl, N, [&](const i64 i, const bool valid) {
    if (!valid || (!plist && !rb[rep[i]])) return;
    const i32 pos = cstart[cell_of[i]] + rank[i];
    tmp[pos] = (i32)i;
    keys[pos] = orig[i];
  });
}

// In-cell order by original particle id (deterministic regardless of the
// atomic ranks): the new position of a member is the number of members of
// its cell with a smaller id, counted over the cell's contiguous keys.
__global__ void k_cellsort(const i32* tmp, const i32* keys, const i32* cell_of, const i32* cstart, i64 N, i32* perm,
                           const DpCtl* ctl, const i32* rep, const i32* rb, const int2* plist) {
  if (!ctl->rebuild) return;
  dp_for_rebuild(plist, ctl, N, [&](const i64 q, const bool valid) {
    if (!valid || (!plist && !rb[rep[q]])) return;  // replicas own fixed slot ranges
    const i32 p = tmp[q];
    const i32 c = cell_of[p];
    const i32 lo = cstart[c], hi = cstart[c + 1];
    const i32 key = keys[q];
    // the cell's keys, four per 16-byte load where aligned (single-cell
    // replicas hold hundreds of particles in one cell)
    i32 rnk = 0;
    i32 k = lo;
    for (; k < hi && (k & 3) != 0; ++k) rnk += (keys[k] < key) ? 1 : 0;
    for (; k + 4 <= hi; k += 4) {
      const int4 v = *(const int4*)(keys + k);
      rnk += ((v.x < key) ? 1 : 0) + ((v.y < key) ? 1 : 0) + ((v.z < key) ? 1 : 0) + ((v.w < key) ? 1 : 0);
    }
    for (; k < hi; ++k) rnk += (keys[k] < key) ? 1 : 0;
    perm[lo + rnk] = p;
  });
}

// Gather every per-particle array into the new order (into scratch), then
// copy back.  Scratch layout: x | v | beta (N*D doubles each), mass (N),
// img (N*D i32), type | orig | cell (N i32 each).
template <int D>
__global__ void k_permute(const i32* perm, i64 N, const double* x, const double* v, const double* beta,
                          const double* mass, const i32* img, const i32* type, const i32* orig,
                          const i32* cell_of, char* scratch, const DpCtl* ctl, const i32* rep, const i32* rb,
                          const int2* plist) {
  if (!ctl->rebuild) return;
  double* sx = (double*)scratch;
  double* sv = sx + N * D;
  double* sb = sv + N * D;
  double* sm = sb + N * D;
  i32* si = (i32*)(sm + N);
  i32* st = si + N * D;
  i32* so = st + N;
  i32* sc = so + N;
  dp_for_rebuild(plist, ctl, N, [&](const i64 q, const bool valid) {
    if (!valid || (!plist && !rb[rep[q]])) return;
    const i64 p = perm[q];
    double t[D];
    dp_ldD<D>(x + p * D, t);
    dp_stD<D>(sx + q * D, t);
    dp_ldD<D>(v + p * D, t);
    dp_stD<D>(sv + q * D, t);
    dp_ldD<D>(beta + p * D, t);
    dp_stD<D>(sb + q * D, t);
#pragma unroll
    for (int d = 0; d < D; ++d) si[q * D + d] = img[p * D + d];
    sm[q] = mass[p];
    st[q] = type[p];
    so[q] = orig[p];
    sc[q] = cell_of[p];
  });
}

// Fixed-point cell-relative coordinate of a position component: the offset
// from the centre of the particle's cell (minimum image), scaled by 2^qbits.
__device__ __forceinline__ i32 dp_quantise(double xd, double cen, double L, double halfL, double qinv) {
  return (i32)rint(dp_minimage_gt(xd - cen, L, halfL) * qinv);
}

// Copy back from scratch, record the reference positions (Simulation.py:381
// positions_neighbour_list = positions.copy()), pack cell coordinates, and
// write the mixed-precision records (fixed-point offsets from the cell
// centre; type in the slot after the coordinates).
template <int D>
__global__ void k_unpermute_store(i64 N, double* x, double* v, double* beta, double* mass, i32* img, i32* type,
                                  i32* orig, i32* cell_of, const char* scratch, double* xref, i32* cellpack,
                                  i32* xq, int rs, double qinv, const i32* rep, const DpRep* reps,
                                  const DpCtl* ctl, const i32* rb, const int2* plist) {
  if (!ctl->rebuild) return;
  const double* sx = (const double*)scratch;
  const double* sv = sx + N * D;
  const double* sb = sv + N * D;
  const double* sm = sb + N * D;
  const i32* si = (const i32*)(sm + N);
  const i32* st = si + N * D;
  const i32* so = st + N;
  const i32* sc = so + N;
  dp_for_rebuild(plist, ctl, N, [&](const i64 q, const bool valid) {
    if (!valid || (!plist && !rb[rep[q]])) return;
    const DpRep& R = reps[rep[q]];
    const i32 cid = sc[q];
    int c[D];
    dp_unlinear<D>((i64)cid - R.c0, R.nc, R.order, c);
    u32 pack = 0;
    double xs[D], t[D];
    dp_ldD<D>(sx + q * D, xs);
    dp_stD<D>(x + q * D, xs);
    dp_stD<D>(xref + q * D, xs);
    dp_ldD<D>(sv + q * D, t);
    dp_stD<D>(v + q * D, t);
    dp_ldD<D>(sb + q * D, t);
    dp_stD<D>(beta + q * D, t);
#pragma unroll
    for (int d = 0; d < D; ++d) {
      pack |= ((u32)c[d]) << (DpPack<D>::BITS * d);
      img[q * D + d] = si[q * D + d];
      if (xq) xq[q * rs + d] = dp_quantise(xs[d], ((double)c[d] + 0.5) * R.w[d], R.L[d], R.halfL[d], qinv);
    }
    mass[q] = sm[q];
    type[q] = st[q];
    orig[q] = so[q];
    cell_of[q] = cid;
    cellpack[q] = (i32)pack;
    if (xq) {
      if (D < rs) xq[q * rs + D] = st[q];
      for (int k = D + 1; k < rs; ++k) xq[q * rs + k] = 0;
    }
  });
}

// Mixed records from the current positions (after host uploads).
template <int D>
__global__ void k_make_records(i64 N, const double* x, const i32* type, const i32* cellpack, i32* xq, int rs,
                               double qinv, const i32* rep, const DpRep* reps) {
  DP_FOR(q, N) {
    const DpRep& R = reps[rep[q]];
    const u32 pack = (u32)cellpack[q];
#pragma unroll
    for (int d = 0; d < D; ++d) {
      const u32 c = (pack >> (DpPack<D>::BITS * d)) & DpPack<D>::MASK;
      xq[q * rs + d] = dp_quantise(x[q * D + d], ((double)c + 0.5) * R.w[d], R.L[d], R.halfL[d], qinv);
    }
    if (D < rs) xq[q * rs + D] = type[q];
    for (int k = D + 1; k < rs; ++k) xq[q * rs + k] = 0;
  }
}

// ---------------------------------------------------------------------------
// Neighbour-list emit, one workgroup per tile.  The halo box's positions at
// the last build (xref) are staged in LDS as fixed-point records in the
// tile's frame (offsets from the cell centre plus the halo cell's offset, as
// in the pair kernels).  A candidate's exact squared distance in fixed-point
// units decides it unless it lies within a few units of the list cutoff;
// those are decided by the FP64 test of neighbor_list.py:24-33 on xref
// (evaluated as s < threshold, exactly sqrt(s) < cutoff).  Entries keep the
// stencil order of neighbor_list.py:150-157 (order 0) / :82-86 (order 1) and
// ascending original id inside a cell; an entry is the candidate's index in
// the tile's staged halo: cell occupancies are fixed between builds, so the
// pair kernels read a neighbour's record straight from LDS with it.  Reading
// xref makes the emit repeatable: a new tile plan re-emits the same list.
// ---------------------------------------------------------------------------
template <int D>
__device__ __forceinline__ int4 dp_emit_rec(const double* xref, const i32* cellpack, i64 slot, const DpRep& P,
                                           const DpTileSmem<D>& S, int u, double qinv) {
  const u32 pack = (u32)cellpack[slot];
  i32 q[4] = {0, 0, 0, 0};
  double xv[D];
  dp_ldD<D>(xref + slot * D, xv);
#pragma unroll
  for (int d = 0; d < D; ++d) {
    const u32 c = (pack >> (DpPack<D>::BITS * d)) & DpPack<D>::MASK;
    const i32 v = dp_quantise(xv[d], ((double)c + 0.5) * P.w[d], P.L[d], P.halfL[d], qinv);
    q[d] = (i32)((u32)v + dp_halo_shift<D>(S, P, u, d));
  }
  return make_int4(q[0], q[1], q[2], q[3]);
}

// Candidate-to-particle differences along each dimension (the minimum image
// along single-cell dimensions when CM), the FP32 screen of their squared sum
// and, for the few candidates the screen cannot decide, the exact one.
template <int D, bool CM>
__device__ __forceinline__ void dp_emit_dd(const i32* qi, const int4& qj, const DpRep& P, i32* dd) {
  const i32 bv[4] = {qj.x, qj.y, qj.z, qj.w};
#pragma unroll
  for (int d = 0; d < D; ++d) {
    i32 t = (i32)((u32)qi[d] - (u32)bv[d]);
    if (CM && P.nc[d] == 1) {
      const i32 h = P.Lq[d] >> 1;
      t = (t > h) ? t - P.Lq[d] : ((t < -h) ? t + P.Lq[d] : t);
    }
    dd[d] = t;
  }
}

// High 16 bits of two fixed-point coordinates packed into one word (x in the
// low half): the operands of the 16-bit screen.
typedef short dp_s16x2 __attribute__((ext_vector_type(2)));
__device__ __forceinline__ u32 dp_pack_hi16(i32 a, i32 b) {
  return ((u32)a >> 16) | ((u32)b & 0xffff0000u);
}

// 16-bit screen of a candidate: sum over dimensions of ((a_d >> 16) -
// (b_d >> 16))^2, each difference wrapping in 16 bits (exact while the
// displacement is below 2^31 units) and the sum saturating at 2^31 - 1.
template <int D>
__device__ __forceinline__ int dp_emit_s16(u32 pxy, u32 pzw, u32 cxy, u32 czw) {
  const dp_s16x2 dxy = __builtin_bit_cast(dp_s16x2, pxy) - __builtin_bit_cast(dp_s16x2, cxy);
  int s = __builtin_amdgcn_sdot2(dxy, dxy, 0, true);
  if (D >= 3) {
    const dp_s16x2 dzw = __builtin_bit_cast(dp_s16x2, pzw) - __builtin_bit_cast(dp_s16x2, czw);
    s = __builtin_amdgcn_sdot2(dzw, dzw, s, true);
  }
  return s;
}

template <int D>
__device__ __forceinline__ float dp_emit_sf(const i32* dd) {
  float s = 0.0f;
#pragma unroll
  for (int d = 0; d < D; ++d) {
    const float t = (float)dd[d];
    s = (d == 0) ? t * t : __builtin_fmaf(t, t, s);
  }
  return s;
}

// Exact squared fixed-point distance: |dd| < 2^31 per dimension (the
// fixed-point scale is chosen so), so a sum of D <= 4 squares fits in 64 bits.
template <int D>
__device__ __forceinline__ u64 dp_emit_sq(const i32* dd) {
  u64 s = 0;
#pragma unroll
  for (int d = 0; d < D; ++d) {
    const u32 ad = (u32)(dd[d] < 0 ? -dd[d] : dd[d]);
    s += (u64)ad * (u64)ad;
  }
  return s;
}

// Emit of one tile's rows.  Every particle of a home cell has the same
// candidates (its stencil rows, in order).  Wavefront w takes particles w,
// w + nwaves, ... of each home cell, G at a time; for each chunk of
// 64 candidates (one per lane) it finds their staged indices and records
// once, then tests the group's particles against the chunk one after the
// other (independent chains; the particles' records and running row lengths
// are wavefront-uniform and live in scalar registers).  A candidate is listed
// when the FP32 screen puts it below the band; inside the screen's slack the
// exact fixed-point distance decides, and inside the fixed-point band the
// reference's FP64 test on xref.
// Rows are padded with zero entries (staged index 0, a valid record of any
// non-empty tile) from their length c up to the next multiple of the pair
// kernels' chunk (four 16-bit or two 32-bit entries, one 8-byte word): the
// pair kernels then read every entry of a row's last chunk without a select
// (entries past the length are excluded from the sums by their position).
// Rows are K entries at most; the padding stays within the stride (K
// rounded up to 8).
template <typename E>
__device__ __forceinline__ u32 dp_row_pad(u32 c) {
  constexpr u32 EPW = 8 / sizeof(E);
  return (EPW - c % EPW) % EPW;
}

// Entry format (efmt = esh | sent << 8): an entry is a neighbour's staged-halo
// index shifted left by esh (mixed precision: the byte offset of its record
// in the pair kernels' LDS record area); with sent, rows are padded with the
// entry of the tile's sentinel record (staged index nrec, after the halo),
// which lies beyond the cut of every home particle, otherwise with 0.
__device__ __forceinline__ int dp_efmt_esh(int efmt) { return efmt & 31; }
__device__ __forceinline__ u32 dp_efmt_pad(int efmt, int nrec) {
  return (efmt >> 8) ? ((u32)nrec << (efmt & 31)) : 0u;
}

template <int D> struct DpEmitG { static constexpr int value = D >= 4 ? 4 : 8; };
template <int D, typename E, bool LDS, bool CM>
__device__ __forceinline__ void dp_emit_particles(const DpTileSmem<D>& S, const DpRep& P, const double* xref,
                                                  const i32* cellpack, double qinv, const int4* srq, int K, E* nl,
                                                  i32* nlcnt, DpCtl* ctl, int efmt, int& st_count) {
  constexpr int N3 = DpPow3<D>::value;
  constexpr int NROW = N3 / 3;  // <= 27 for D <= 4: one row per lane
  const int esh = dp_efmt_esh(efmt);
  const u32 pv = dp_efmt_pad(efmt, S.nrec);
  constexpr int G = DpEmitG<D>::value;  // particles per candidate pass
  const int tid = threadIdx.x;
  const int lane = tid & 63;
  const int wave = tid >> 6;
  const int nwaves = blockDim.x >> 6;
  const i64 Ks = (i64)((K + 7) & ~7);  // row stride: rows are read 8 bytes at a time
  const u64 lo = P.nlq_lo, hi = P.nlq_hi;
  const float lo_f = P.nlq_lo_f, hi_f = P.nlq_hi_f;
  const int lo16 = P.nlq16_lo, hi16 = P.nlq16_hi;
  for (int m = 0; m < P.nhome; ++m) {
    const int uh = S.home_u[m];
    if (uh < 0) continue;
    // this tile's slice of the cell's particles
    const int ncell = S.hbase[uh + 1] - S.hbase[uh];
    const int jlo = min(max(S.k0 - S.home_p[m], 0), ncell);
    const int jhi = min(max(S.k1 - S.home_p[m], 0), ncell);
    const int self0 = S.hbase[uh] + jlo;
    const int np = jhi - jlo;
    const i64 slot0 = S.hstart[uh] + jlo;
    // lane r < NROW holds stencil row r: its valid cells are consecutive halo cells
    int r_lo = 0, r_len = 0, r_u0 = 0;
    if (lane < NROW) {
      bool any = false;
      int hcell = 0;
      for (int dx = 0; dx < 3; ++dx) {
        const int st = 3 * lane + dx;
        int sd[D];
        dp_stencil_digits<D>(st, P.order, sd);
        bool ok = true;
#pragma unroll
        for (int d = 0; d < D; ++d)
          if (P.nc[d] == 1 && sd[d] != 0) ok = false;
        if (!ok) continue;
        const int u = uh + S.st_off[st];
        if (!any) { any = true; r_u0 = u; r_lo = S.hbase[u]; }
        hcell = S.hbase[u + 1];
      }
      r_len = any ? hcell - r_lo : 0;
    }
    int incl = r_len;
#pragma unroll
    for (int off = 1; off < 64; off <<= 1) {
      const int vv = __shfl_up(incl, off);
      if (lane >= off) incl += vv;
    }
    const int r_pref = incl - r_len;
    const int ncand = __shfl(incl, NROW - 1);
    // this wavefront's particles: k = wave + nwaves * j, j = 0 .. npw - 1
    const int npw = (np > wave) ? (np - wave + nwaves - 1) / nwaves : 0;
    for (int j0 = 0; j0 < npw; j0 += G) {
      const int nj = min(G, npw - j0);
      i32 qi[G][4];
      u32 pxy[G], pzw[G];
      int selfv[G];
      u32 cnt[G];
#pragma unroll
      for (int g = 0; g < G; ++g) {
        const int k = wave + nwaves * (j0 + min(g, nj - 1));
        selfv[g] = self0 + k;
        const int4 r = LDS ? srq[self0 + k] : dp_emit_rec<D>(xref, cellpack, slot0 + k, P, S, uh, qinv);
        qi[g][0] = __builtin_amdgcn_readfirstlane(r.x);
        qi[g][1] = __builtin_amdgcn_readfirstlane(r.y);
        qi[g][2] = __builtin_amdgcn_readfirstlane(r.z);
        qi[g][3] = __builtin_amdgcn_readfirstlane(r.w);
        pxy[g] = dp_pack_hi16(qi[g][0], qi[g][1]);
        pzw[g] = dp_pack_hi16(qi[g][2], qi[g][3]);
        cnt[g] = 0;
      }
      for (int base = 0; base < ncand; base += 64) {
        const int c = base + lane;
        // row of candidate c: largest r with r_pref[r] <= c (fixed-step search)
        int row = 0;
#pragma unroll
        for (int step = 16; step >= 1; step >>= 1) {
          const int cand = row + step;
          const int pm = __shfl(r_pref, cand < NROW ? cand : NROW - 1);
          if (cand < NROW && pm <= c) row = cand;
        }
        const int lo_row = __shfl(r_lo, row);
        const int pf_row = __shfl(r_pref, row);
        const int u0_row = __shfl(r_u0, row);
        const bool valid = c < ncand;
        const int idx = valid ? lo_row + (c - pf_row) : 0;
        // halo cell of the candidate: found here for global records, in the
        // (rare) exact path for staged ones
        int uj = u0_row;
        int4 qj = make_int4(0, 0, 0, 0);
        if (valid) {
          if (LDS) {
            qj = srq[idx];
          } else {
            while (S.hbase[uj + 1] <= idx) ++uj;
            qj = dp_emit_rec<D>(xref, cellpack, (i64)S.hstart[uj] + (idx - S.hbase[uj]), P, S, uj, qinv);
          }
        }
        // screen: bit g of `hit` / `band` for particle g of the group
        u32 hit = 0u, band = 0u;
        if (!CM) {
          // 16-bit screen: two packed subtractions and two dot products per
          // particle and candidate
          const u32 cxy = dp_pack_hi16(qj.x, qj.y), czw = dp_pack_hi16(qj.z, qj.w);
#pragma unroll
          for (int g = 0; g < G; ++g) {
            if (g < nj) {
              const int s16 = dp_emit_s16<D>(pxy[g], pzw[g], cxy, czw);
              const bool cand_ok = valid && idx != selfv[g];
              hit |= (cand_ok && s16 < lo16) ? (1u << g) : 0u;
              band |= (cand_ok && s16 >= lo16 && s16 <= hi16) ? (1u << g) : 0u;
            }
          }
        } else {
#pragma unroll
          for (int g = 0; g < G; ++g) {
            if (g < nj) {
              i32 dd[D];
              dp_emit_dd<D, CM>(qi[g], qj, P, dd);
              const float sf = dp_emit_sf<D>(dd);
              const bool cand_ok = valid && idx != selfv[g];
              hit |= (cand_ok && sf < lo_f) ? (1u << g) : 0u;
              band |= (cand_ok && !(sf < lo_f) && sf <= hi_f) ? (1u << g) : 0u;
            }
          }
        }
        if (__ballot(band != 0u) != 0ull) {
          // the screen's slack: exact fixed-point distance, then within a few
          // units of the cutoff the reference's test on xref
          if (LDS && band != 0u)
            while (S.hbase[uj + 1] <= idx) ++uj;
          for (int g = 0; g < nj; ++g) {
            if (band & (1u << g)) {
              const int k = wave + nwaves * (j0 + g);
              const int4 r = LDS ? srq[self0 + k] : dp_emit_rec<D>(xref, cellpack, slot0 + k, P, S, uh, qinv);
              const i32 qg[4] = {r.x, r.y, r.z, r.w};
              i32 dd[D];
              dp_emit_dd<D, CM>(qg, qj, P, dd);
              const u64 sq = dp_emit_sq<D>(dd);
              bool h = sq < lo;
              if (!h && sq <= hi) {
                const i64 slot_i = slot0 + k;
                const i64 slot_j = (i64)S.hstart[uj] + (idx - S.hbase[uj]);
                double ss = 0.0;
#pragma unroll
                for (int d = 0; d < D; ++d) {
                  const double tt =
                      dp_minimage_lt(dp_sub(xref[slot_i * D + d], xref[slot_j * D + d]), P.L[d], P.halfL[d]);
                  ss = (d == 0) ? dp_mul(tt, tt) : dp_add(ss, dp_mul(tt, tt));
                }
                h = ss < P.nl_thresh;
              }
              if (h) hit |= 1u << g;
            }
          }
        }
        // rows: the hits of each particle in lane order after its running length
#pragma unroll
        for (int g = 0; g < G; ++g) {
          if (g < nj) {
            const bool h = (hit >> g) & 1u;
            const u64 mask = __ballot(h);
            const u32 below = __builtin_amdgcn_mbcnt_hi((u32)(mask >> 32), __builtin_amdgcn_mbcnt_lo((u32)mask, 0u));
            const u32 pos = cnt[g] + below;
            if (h && pos < (u32)K) nl[(slot0 + (selfv[g] - self0)) * Ks + pos] = (E)((u32)idx << esh);
            cnt[g] += (u32)__popcll(mask);
          }
        }
      }
#pragma unroll
      for (int g = 0; g < G; ++g) {
        if (g < nj) {
          // padding entries up to the pair kernels' chunk of the row's end
          const u32 c0 = min(cnt[g], (u32)K);
          if ((u32)lane < dp_row_pad<E>(c0)) nl[(slot0 + (selfv[g] - self0)) * Ks + c0 + lane] = (E)pv;
          if (lane == 0) {
            nlcnt[slot0 + (selfv[g] - self0)] = (i32)cnt[g];
            if (cnt[g] > (u32)K) ctl->overflow = 1;
          }
          st_count = max(st_count, (int)cnt[g]);
        }
      }
    }
  }
}

// Exact decision for a candidate inside the 16-bit screen's slack (rare; kept
// out of line so that its operands do not hold registers in the screen loop):
// the exact fixed-point distance, then within a few units of the cutoff the
// reference's test on xref.  si / sj: staged indices; slot_i: the particle's slot.
template <int D>
__device__ __attribute__((noinline)) bool dp_emit_exact(const DpTileSmem<D>& S, const DpRep& P, const double* xref,
                                                       const int4* srq, int si, int sj, i64 slot_i) {
  const int4 r = srq[si];
  const i32 qg[4] = {r.x, r.y, r.z, r.w};
  i32 dd[D];
  dp_emit_dd<D, false>(qg, srq[sj], P, dd);
  const u64 sq = dp_emit_sq<D>(dd);
  if (sq < P.nlq_lo) return true;
  if (sq > P.nlq_hi) return false;
  const int uj = dp_halo_of<D>(S, P.nhalo, sj);
  const i64 slot_j = (i64)S.hstart[uj] + (sj - S.hbase[uj]);
  double ss = 0.0;
#pragma unroll
  for (int d = 0; d < D; ++d) {
    const double tt = dp_minimage_lt(dp_sub(xref[slot_i * D + d], xref[slot_j * D + d]), P.L[d], P.halfL[d]);
    ss = (d == 0) ? dp_mul(tt, tt) : dp_add(ss, dp_mul(tt, tt));
  }
  return ss < P.nl_thresh;
}

// One home cell's table pass: its tables and wavefront-uniform values.
template <int D, typename E> struct DpEmitCell {
  const DpTileSmem<D>& S;
  const DpRep& P;
  const double* xref;
  const int4* srq;
  const u16* ctab;
  E* nl;
  i32* nlcnt;
  DpCtl* ctl;
  int K;
  i64 Ks;
  int self0;
  i64 slot0;
  int ncand, cs0, lo16, hi16;
  int esh;  // entry = staged index << esh
  u32 pv;   // padding entry
};

// Rows of NJ particles k0, k0 + NW, .. of one home cell against its candidate
// table; returns the longest row.  Per chunk of 64 candidates and particle:
// the 16-bit screen (lane masks of single compares, combined in scalar code),
// then the hits' stores; a candidate in the screen's slack (rare) sends that
// particle's chunk through the exact decision first.  A hit's store address
// is a 32-bit byte offset from the group's first row, clamped to the row's
// last entry (a row longer than K is reported as an overflow; its entries are
// not used).
template <int D, typename E, int NJ>
__device__ __forceinline__ int dp_emit_group(const DpEmitCell<D, E>& C, int k0) {
  constexpr int NW = DP_TILE_THREADS / 64;
  constexpr int ESH = sizeof(E) == 2 ? 1 : 2;
  const int lane = threadIdx.x & 63;
  u32 pxy[NJ], pzw[NJ], cnt[NJ];
#pragma unroll
  for (int g = 0; g < NJ; ++g) {
    const int4 r = C.srq[C.self0 + k0 + NW * g];
    pxy[g] = __builtin_amdgcn_readfirstlane(dp_pack_hi16(r.x, r.y));
    pzw[g] = __builtin_amdgcn_readfirstlane(dp_pack_hi16(r.z, r.w));
    cnt[g] = 0;
  }
  // the group's rows lie within NJ * NW rows of the first one
  char* const rows = (char*)(C.nl + (C.slot0 + k0) * C.Ks);
  const u32 rstepb = (u32)(NW * C.Ks) << ESH;
  // the clamp bound in a vector register: scalar registers are the scarce
  // resource of this loop (the bound would be reloaded from a spill slot)
  u32 last;
  asm volatile("v_mov_b32 %0, %1" : "=v"(last) : "s"((u32)C.Ks - 1u));
  const int self_c = C.cs0 + k0;  // candidate position of particle g's self entry: self_c + NW * g
  for (int base = 0; base < C.ncand; base += 64) {
    const int c = base + lane;
    const bool valid = c < C.ncand;
    const int idx = valid ? (int)C.ctab[c] : C.self0;
    const int4 qj = C.srq[idx];
    const u32 cxy = dp_pack_hi16(qj.x, qj.y), czw = dp_pack_hi16(qj.z, qj.w);
    const u64 vmask = __builtin_amdgcn_ballot_w64(valid);
    const E ev = (E)((u32)idx << C.esh);
#pragma unroll
    for (int g = 0; g < NJ; ++g) {
      // lane masks from single compares, combined in scalar code
      const int s16 = dp_emit_s16<D>(pxy[g], pzw[g], cxy, czw);
      const u64 ltm = __builtin_amdgcn_ballot_w64(s16 < C.lo16);
      const u64 lem = __builtin_amdgcn_ballot_w64(s16 <= C.hi16);
      const u64 okm = vmask & __builtin_amdgcn_ballot_w64(c != self_c + NW * g);
      u64 hm = ltm & okm;
      const u64 slack = lem & ~ltm & okm;
      if (slack != 0ull) {
        // candidates in the screen's slack (rare): the exact decision
        bool x = false;
        if (__builtin_amdgcn_inverse_ballot_w64(slack))
          x = dp_emit_exact<D>(C.S, C.P, C.xref, C.srq, C.self0 + k0 + NW * g, idx, C.slot0 + k0 + NW * g);
        hm |= __builtin_amdgcn_ballot_w64(x);
      }
      if (__builtin_amdgcn_inverse_ballot_w64(hm)) {
        // the running length enters as mbcnt's addend
        const u32 e = min(__builtin_amdgcn_mbcnt_hi((u32)(hm >> 32), __builtin_amdgcn_mbcnt_lo((u32)hm, cnt[g])),
                          last);
        *(E*)(rows + (u64)((e << ESH) + (u32)g * rstepb)) = ev;
      }
      cnt[g] += (u32)__popcll(hm);
    }
  }
  int mx = 0;
#pragma unroll
  for (int g = 0; g < NJ; ++g) {
    // padding entries up to the pair kernels' chunk of the row's end
    const u32 c0 = min(cnt[g], (u32)C.K);
    if ((u32)lane < dp_row_pad<E>(c0)) *(E*)(rows + (u64)(((c0 + (u32)lane) << ESH) + (u32)g * rstepb)) = (E)C.pv;
    if (lane == 0) {
      C.nlcnt[C.slot0 + k0 + NW * g] = (i32)cnt[g];
      if (cnt[g] > (u32)C.K) C.ctl->overflow = 1;
    }
    mx = max(mx, (int)cnt[g]);
  }
  return mx;
}

// A home cell's box in its tile's fixed-point frame, in the high 16 bits of
// the coordinates (the 16-bit screen's units): centre, half width widened
// so that the gap computed below is a lower bound of the true gap / 2^16
// (each high-half difference is within 1 of the true one / 2^16), and the
// squared distance in those units beyond which no pair is listed.
typedef unsigned short dp_u16x2 __attribute__((ext_vector_type(2)));
struct DpEmitBox {
  u32 cxy, czw;       // centre (the cell's shift in the tile frame), high halves packed (x, y) and (z, w)
  dp_u16x2 hxy, hzw;  // half widths + slack, same packing
  u32 thr;            // P.nlq_hi / 2^32
};
template <int D>
__device__ __forceinline__ DpEmitBox dp_emit_box(const DpTileSmem<D>& S, const DpRep& P, int uh) {
  u32 c[4] = {0u, 0u, 0u, 0u};
  unsigned short h[4] = {0, 0, 0, 0};
#pragma unroll
  for (int d = 0; d < D; ++d) {
    c[d] = __builtin_amdgcn_readfirstlane(dp_halo_shift<D>(S, P, uh, d));
    h[d] = (unsigned short)min((((u32)(P.Wq[d] / 2) + 4u + 0xffffu) >> 16) + 2u, 0xffffu);
  }
  DpEmitBox b;
  b.cxy = dp_pack_hi16((i32)c[0], (i32)c[1]);
  b.czw = dp_pack_hi16((i32)c[2], (i32)c[3]);
  b.hxy = dp_u16x2{h[0], h[1]};
  b.hzw = dp_u16x2{h[2], h[3]};
  const u64 t = P.nlq_hi >> 32;
  b.thr = t > 0xffffffffull ? 0xffffffffu : (u32)t;
  return b;
}
// Whether a staged record can be within the list cutoff of a point of the
// box: per dimension the gap max(|high-half difference| - half width, 0)
// (packed 16-bit operations), their squares summed with saturation.  A gap is
// at most the true gap / 2^16 and a saturated sum is smaller than the true
// one, so a record is dropped only when every pair with the box's particles
// lies beyond P.nlq_hi (the fixed-point bound above which no pair is listed).
template <int D>
__device__ __forceinline__ bool dp_emit_in_box(const int4 q, const DpEmitBox& b) {
  const dp_s16x2 dxy = __builtin_bit_cast(dp_s16x2, dp_pack_hi16(q.x, q.y)) - __builtin_bit_cast(dp_s16x2, b.cxy);
  const dp_u16x2 gxy = __builtin_elementwise_sub_sat(__builtin_bit_cast(dp_u16x2, __builtin_elementwise_max(dxy, -dxy)),
                                                     b.hxy);
  u32 s = __builtin_amdgcn_udot2(gxy, gxy, 0u, true);
  if (D >= 3) {
    const dp_s16x2 dzw = __builtin_bit_cast(dp_s16x2, dp_pack_hi16(q.z, q.w)) - __builtin_bit_cast(dp_s16x2, b.czw);
    const dp_u16x2 gzw = __builtin_elementwise_sub_sat(
        __builtin_bit_cast(dp_u16x2, __builtin_elementwise_max(dzw, -dzw)), b.hzw);
    s = __builtin_amdgcn_udot2(gzw, gzw, s, true);
  }
  return s <= b.thr;
}

// Emit of one tile's rows (up to 8 particles per candidate pass) for tiles staged in LDS with no single-cell
// dimension.  Same rows, in the same order, as dp_emit_particles.  Per home
// cell, the workgroup first writes the candidate list (staged indices, in
// stencil-row order, without the records that the home cell's box puts
// beyond the list cutoff of all its particles) to an LDS table `ctab`, so
// that a lane finds its candidate with one LDS read.  The group's particle index, row pointer and
// running row length are wavefront-uniform (scalar registers); the screen's
// results stay lane masks: a hit is stored under its own compare, at
// mbcnt(mask) after the running length.  The self entry sits at a known
// candidate position (the home cell is the centre of its stencil row).
#define DP_EMIT_NROW 27  // stencil rows of D <= 4
template <int D, typename E>
__device__ __forceinline__ void dp_emit_particles_tab(const DpTileSmem<D>& S, const DpRep& P, const double* xref,
                                                      const int4* srq, u16* ctab, i32* ekeep, int K, E* nl,
                                                      i32* nlcnt, DpCtl* ctl, int efmt, int& st_count) {
  constexpr int N3 = DpPow3<D>::value;
  constexpr int NROW = N3 / 3;
  constexpr int G = 8;
  constexpr int NW = DP_TILE_THREADS / 64;
  constexpr int RC = (N3 / 2) / 3;  // stencil row holding the centre cell (its middle cell)
  const int tid = threadIdx.x;
  const int lane = tid & 63;
  const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);
  const i64 Ks = (i64)((K + 7) & ~7);
  const int lo16 = P.nlq16_lo, hi16 = P.nlq16_hi;
  for (int m = 0; m < P.nhome; ++m) {
    // (LDS values that are the same in every lane: read into scalar registers)
    const int uh = __builtin_amdgcn_readfirstlane(S.home_u[m]);
    if (uh < 0) continue;  // uniform over the workgroup: no barrier is skipped by part of it
    // this tile's slice of the cell's particles
    const int ncell = __builtin_amdgcn_readfirstlane(S.hbase[uh + 1]) - __builtin_amdgcn_readfirstlane(S.hbase[uh]);
    const int hp = __builtin_amdgcn_readfirstlane(S.home_p[m]);
    const int jlo = min(max(S.k0 - hp, 0), ncell);
    const int jhi = min(max(S.k1 - hp, 0), ncell);
    const int self0 = __builtin_amdgcn_readfirstlane(S.hbase[uh]) + jlo;
    const int np = jhi - jlo;
    const i64 slot0 = __builtin_amdgcn_readfirstlane(S.hstart[uh]) + jlo;
    // lane r < NROW: stencil row r = cells 3r .. 3r + 2, consecutive halo cells
    int r_lo = 0, r_len = 0;
    if (lane < NROW) {
      const int u0 = uh + S.st_off[3 * lane];
      r_lo = S.hbase[u0];
      r_len = S.hbase[uh + S.st_off[3 * lane + 2] + 1] - r_lo;
    }
    // candidates: the stencil rows' records that can lie within the list
    // cutoff of some particle of the home cell (the box test), in
    // stencil-row order; with cells 1.1 list cutoffs wide, ~62 % of the
    // stencil in 3-D.  Pass 1 counts each row's (wavefront w: rows w,
    // w + NW, ..; in the home cell's row also those before the home cell,
    // whose own particles are all kept)
    const DpEmitBox box = dp_emit_box<D>(S, P, uh);
    const int hb = __builtin_amdgcn_readfirstlane(S.hbase[uh]);
    for (int r = wave; r < NROW; r += NW) {
      const int rl = __builtin_amdgcn_readlane(r_lo, r);
      const int rn = __builtin_amdgcn_readlane(r_len, r);
      int kept = 0, before = 0;
      for (int i0 = 0; i0 < rn; i0 += 64) {
        const int i = i0 + lane;
        const bool k = i < rn && dp_emit_in_box<D>(srq[rl + i], box);
        kept += __popcll(__builtin_amdgcn_ballot_w64(k));
        if (r == RC) before += __popcll(__builtin_amdgcn_ballot_w64(k && rl + i < hb));
      }
      if (lane == 0) {
        ekeep[r] = kept;
        if (r == RC) ekeep[DP_EMIT_NROW] = before;
      }
    }
    __syncthreads();
    int incl = (lane < NROW) ? ekeep[lane] : 0;
    const int r_keep = incl;
#pragma unroll
    for (int off = 1; off < 64; off <<= 1) {
      const int vv = __shfl_up(incl, off);
      if (lane >= off) incl += vv;
    }
    const int r_pref = incl - r_keep;
    const int ncand = __builtin_amdgcn_readlane(incl, NROW - 1);
    // candidate position of the home cell's first particle
    const int cs0 = __builtin_amdgcn_readlane(r_pref, RC) + ekeep[DP_EMIT_NROW] + (self0 - hb);
    // pass 2, the table: row r's kept staged indices at r_pref[r] ..
    for (int r = wave; r < NROW; r += NW) {
      const int rl = __builtin_amdgcn_readlane(r_lo, r);
      const int rn = __builtin_amdgcn_readlane(r_len, r);
      int pos = __builtin_amdgcn_readlane(r_pref, r);
      for (int i0 = 0; i0 < rn; i0 += 64) {
        const int i = i0 + lane;
        const bool k = i < rn && dp_emit_in_box<D>(srq[rl + i], box);
        const u64 km = __builtin_amdgcn_ballot_w64(k);
        if (k) ctab[__builtin_amdgcn_mbcnt_hi((u32)(km >> 32), __builtin_amdgcn_mbcnt_lo((u32)km, (u32)pos))] =
            (u16)(rl + i);
        pos += __popcll(km);
      }
    }
    __syncthreads();
    // this wavefront's particles: k = wave + NW * j, j = 0 .. npw - 1, in
    // groups of up to G (the group size is a template argument of the pass)
    const int npw = (np > wave) ? (np - wave + NW - 1) / NW : 0;
    for (int j0 = 0; j0 < npw; j0 += G) {
      const int nj = min(G, npw - j0);
      const int k0 = wave + NW * j0;
      const DpEmitCell<D, E> C{S,     P,     xref, srq, ctab, nl,   nlcnt, ctl, K, Ks, self0, slot0, ncand, cs0,
                               lo16,  hi16,  dp_efmt_esh(efmt), dp_efmt_pad(efmt, S.nrec)};
      int c = 0;
      switch (nj) {
        case 1: c = dp_emit_group<D, E, 1>(C, k0); break;
        case 2: c = dp_emit_group<D, E, 2>(C, k0); break;
        case 3: c = dp_emit_group<D, E, 3>(C, k0); break;
        case 4: c = dp_emit_group<D, E, 4>(C, k0); break;
        case 5: c = dp_emit_group<D, E, (G >= 5 ? 5 : 1)>(C, k0); break;
        case 6: c = dp_emit_group<D, E, (G >= 6 ? 6 : 1)>(C, k0); break;
        case 7: c = dp_emit_group<D, E, (G >= 7 ? 7 : 1)>(C, k0); break;
        default: c = dp_emit_group<D, E, (G >= 8 ? 8 : 1)>(C, k0); break;
      }
      st_count = max(st_count, c);
    }
    __syncthreads();  // the next home cell rewrites the table
  }
}

// Emit of one tile's rows with one lane per home particle, for cells of a
// few particles (2-D and 4-D workloads), where the candidate-table pass
// would leave most lanes idle.  Same rows, in the same order, as
// dp_emit_particles: each lane walks its stencil rows in order and appends
// its hits to its own row, four 16-bit (two 32-bit) entries per 8-byte store.
// Lanes of one home cell read the same staged record at the same time.
template <int D, typename E>
__device__ __forceinline__ void dp_emit_lanes(const DpTileSmem<D>& S, const DpRep& P, const double* xref,
                                              const int4* srq, int K, E* nl, i32* nlcnt, DpCtl* ctl, int efmt,
                                              int& st_count) {
  constexpr int N3 = DpPow3<D>::value;
  constexpr int EB = 8 * (int)sizeof(E);    // bits per entry
  constexpr int EPW = 64 / EB;              // entries per 8-byte store
  const int esh = dp_efmt_esh(efmt);
  const u32 pv = dp_efmt_pad(efmt, S.nrec);
  const i64 Ks = (i64)((K + 7) & ~7);
  const int lo16 = P.nlq16_lo, hi16 = P.nlq16_hi;
  for (int k = S.k0 + threadIdx.x; k < S.k1; k += blockDim.x) {
    const int m = dp_home_of<D>(S, P.nhome, k);
    const int uh = S.home_u[m];
    const int loc = k - S.home_p[m];
    const int self = S.hbase[uh] + loc;
    const i64 slot = (i64)S.hstart[uh] + loc;
    const int4 ri = srq[self];
    const u32 pxy = dp_pack_hi16(ri.x, ri.y), pzw = dp_pack_hi16(ri.z, ri.w);
    E* row = nl + slot * Ks;
    u32 cnt = 0;
    u64 pend = 0;  // entries of the current 8-byte group
    // stencil row r = cells 3r .. 3r + 2: consecutive halo cells, one range
    // of staged records
    for (int r = 0; r < N3 / 3; ++r) {
      const int u = uh + S.st_off[3 * r];
      const int j1 = S.hbase[u + 3];
      for (int j = S.hbase[u]; j < j1; ++j) {
        const int4 qj = srq[j];
        const int s16 = dp_emit_s16<D>(pxy, pzw, dp_pack_hi16(qj.x, qj.y), dp_pack_hi16(qj.z, qj.w));
        bool h = s16 < lo16 && j != self;
        // the screen's slack (rare): the exact decision
        if (!h && s16 <= hi16 && j != self) h = dp_emit_exact<D>(S, P, xref, srq, self, j, slot);
        if (h) {
          pend |= (u64)((u32)j << esh) << (EB * (cnt % EPW));
          if (cnt % EPW == EPW - 1) {
            if (cnt - (EPW - 1) < (u32)K) *(u64*)(row + (cnt - (EPW - 1))) = pend;
            pend = 0;
          }
          ++cnt;
        }
      }
    }
    // the last, partial group, its entries past cnt the padding entry (the
    // pad of dp_row_pad)
    const u32 g0 = cnt - cnt % EPW;
    if (cnt % EPW != 0 && g0 < (u32)K) {
      for (u32 q = cnt % EPW; q < (u32)EPW; ++q) pend |= (u64)pv << (EB * q);
      *(u64*)(row + g0) = pend;
    }
    nlcnt[slot] = (i32)cnt;
    if (cnt > (u32)K) ctl->overflow = 1;
    st_count = max(st_count, (int)cnt);
  }
}

// Tiles whose halo fits the LDS record area; the others are marked in their
// table (word 6) and emitted by k_emit_tile_far.  Keeping the global-memory
// path out of this kernel keeps it at <= 128 VGPRs (4 wavefronts per SIMD).
template <int D, typename E>
__global__ void __launch_bounds__(DP_TILE_THREADS, 4) k_emit_tile(const double* __restrict__ xref,
                                                               const i32* __restrict__ cellpack, double qinv,
                                                               const i32* __restrict__ cstart,
                                                               const DpRep* __restrict__ reps, int R,
                                                               i64 ntiles, int K, int cap, E* nl, i32* nlcnt,
                                                               i32* tiletab, int tstride, DpCtl* ctl, int force,
                                                               int lanes, const i32* __restrict__ todo, int efmt,
                                                               int cap_pair) {
  if (!force && !ctl->rebuild) return;
  // non-finite positions were binned into arbitrary cells: skip the (then
  // unbounded) emit; the error is reported after the step block
  if (ctl->nonfinite) return;
  __shared__ DpTileSmem<D> S;
  extern __shared__ __align__(16) int4 srq[];  // cap staged records, then cap u16 candidate indices
  u16* ctab = (u16*)(srq + cap);
  __shared__ i32 ekeep[DP_EMIT_NROW + 1];  // candidates kept per stencil row; those before the home cell
  const int tid = threadIdx.x;
  const int lane = tid & 63;
  const int wave = tid >> 6;
  const int nwaves = blockDim.x >> 6;
  int st_count = 0, st_cell = 0, st_halo = 0;  // running maxima (per wavefront / thread 0)
  // the tiles to emit: all of them, or (batches) the list of the rebuilding
  // replicas' tiles built by k_emit_list
  const i64 ntodo = todo ? (i64)ctl->emit_ntodo : ntiles;
  for (i64 i = blockIdx.x; i < ntodo; i += gridDim.x) {
    const i64 t = todo ? (i64)todo[i] : i;
    dp_tile_setup<D>(S, reps, R, cstart, t, cap);
    const DpRep& P = reps[__builtin_amdgcn_readfirstlane(S.rep)];  // uniform: scalar loads
    const bool fits = S.fits != 0;
    dp_tile_store<D>(S, P, tiletab + t * (i64)tstride);
    // entry format of the tile: byte offsets (efmt) where the halo fits the
    // largest LDS record area the pair kernels may be given (cap_pair: 48 KB
    // of records, so byte offsets stay below 2^16; the area in use may grow
    // after this emit, when a read-back finds larger halos); plain
    // staged-halo indices otherwise
    const int efmt_t = ((efmt & 31) != 0 && S.nrec + (efmt >> 8) <= cap_pair) ? efmt : 0;
    if (tid == 0) {
      tiletab[t * (i64)tstride + 6] = fits ? 0 : 1;
      tiletab[t * (i64)tstride + 7] = efmt_t != 0 ? 1 : 0;
      if (!fits) ctl->far_tiles = 1;  // plain store: k_emit_tile_far has work
    }
    if (fits) {
      for (int idx = tid; idx < S.nrec; idx += blockDim.x) {
        const int u = dp_halo_of<D>(S, P.nhalo, idx);
        srq[idx] = dp_emit_rec<D>(xref, cellpack, (i64)S.hstart[u] + (idx - S.hbase[u]), P, S, u, qinv);
      }
    }
    if (tid == 0) {
      // entries address the staged halo: E must hold the last entry (the
      // sentinel's, nrec << esh, or (nrec - 1) << esh)
      if (sizeof(E) < 4 && S.nrec > dp_efmt_hlim(efmt_t)) ctl->cell_overflow = 1;
      st_cell = max(st_cell, S.maxc);
      st_halo = max(st_halo, S.nrec);
    }
    __syncthreads();
    const bool cm = P.collapsed_mask != 0;
    if (fits) {
      if (cm)
        dp_emit_particles<D, E, true, true>(S, P, xref, cellpack, qinv, srq, K, nl, nlcnt, ctl, efmt_t, st_count);
      else if (lanes) dp_emit_lanes<D, E>(S, P, xref, srq, K, nl, nlcnt, ctl, efmt_t, st_count);
      else dp_emit_particles_tab<D, E>(S, P, xref, srq, ctab, ekeep, K, nl, nlcnt, ctl, efmt_t, st_count);
    }
    __syncthreads();
  }
  // statistics: one update per workgroup, spread over the slots
  __shared__ int red[DP_TILE_THREADS / 64];
  if (lane == 0) red[wave] = st_count;
  __syncthreads();
  if (tid == 0) {
    for (int w = 1; w < nwaves; ++w) st_count = max(st_count, red[w]);
    i32* sl = ctl->slot + (blockIdx.x % DP_STAT_SLOTS) * DP_STAT_STRIDE;
    atomicMax(sl + DP_STAT_COUNT, st_count);
    atomicMax(sl + DP_STAT_CELL, st_cell);
    atomicMax(sl + DP_STAT_HALO, st_halo);
  }
}

// The emit's tile list of a batch: the tiles of the replicas that rebuild
// this step, replica by replica (one workgroup; a block scan of the
// rebuilding replicas' tile counts, then each thread writes its replica's
// tiles).  Outside rebuilds the list is empty.
__global__ void __launch_bounds__(256) k_emit_list(const DpRep* __restrict__ reps, int R, const i32* __restrict__ rb,
                                                   i32* todo, DpCtl* ctl) {
  __shared__ int sh[256];
  __shared__ int carry;
  const int tid = threadIdx.x;
  if (!ctl->rebuild) {
    if (tid == 0) ctl->emit_ntodo = 0;
    return;
  }
  if (tid == 0) carry = 0;
  __syncthreads();
  for (int base = 0; base < R; base += 256) {
    const int r = base + tid;
    const int cnt = (r < R && rb[r]) ? (int)reps[r].ntiles : 0;
    sh[tid] = cnt;
    __syncthreads();
    for (int off = 1; off < 256; off <<= 1) {
      const int v = tid >= off ? sh[tid - off] : 0;
      __syncthreads();
      sh[tid] += v;
      __syncthreads();
    }
    const int first = carry + sh[tid] - cnt;
    if (cnt > 0) {
      const int t0 = (int)reps[r].t0;
      for (int j = 0; j < cnt; ++j) todo[first + j] = t0 + j;
    }
    __syncthreads();
    if (tid == 255) carry += sh[255];
    __syncthreads();
  }
  if (tid == 0) ctl->emit_ntodo = carry;
}

// Batches: the lists that restrict the rebuild pipeline to the replicas that
// rebuild in this step (one workgroup; block scans over the replicas): the
// DP_BLOCK-particle blocks [start, end) of their slot ranges (plist,
// ctl->pblk_n) and the replicas themselves (rlist, ctl->rb_n).  The other
// replicas keep their cells, cell starts (their cell counts stay zero) and
// slots.
__global__ void __launch_bounds__(256) k_rebuild_list(const DpRep* __restrict__ reps, int R,
                                                      const i32* __restrict__ rb, int2* plist, i32* rlist,
                                                      DpCtl* ctl) {
  __shared__ int sb[256], sr[256];
  __shared__ int carry_b, carry_r;
  const int tid = threadIdx.x;
  if (!ctl->rebuild) return;
  if (tid == 0) carry_b = carry_r = 0;
  __syncthreads();
  for (int base = 0; base < R; base += 256) {
    const int r = base + tid;
    const bool on = r < R && rb[r];
    const int nb = on ? (int)((reps[r].n + DP_BLOCK - 1) / DP_BLOCK) : 0;
    sb[tid] = nb;
    sr[tid] = on ? 1 : 0;
    __syncthreads();
    for (int off = 1; off < 256; off <<= 1) {
      const int vb = tid >= off ? sb[tid - off] : 0;
      const int vr = tid >= off ? sr[tid - off] : 0;
      __syncthreads();
      sb[tid] += vb;
      sr[tid] += vr;
      __syncthreads();
    }
    if (on) {
      const int p0 = (int)reps[r].p0, p1 = (int)(reps[r].p0 + reps[r].n);
      const int first = carry_b + sb[tid] - nb;
      for (int k = 0; k < nb; ++k) plist[first + k] = make_int2(p0 + k * DP_BLOCK, p1);
      rlist[carry_r + sr[tid] - 1] = r;
    }
    __syncthreads();
    if (tid == 255) {
      carry_b += sb[255];
      carry_r += sr[255];
    }
    __syncthreads();
  }
  if (tid == 0) {
    ctl->pblk_n = carry_b;
    ctl->rb_n = carry_r;
  }
}

// Batches: the cell-count scan of each rebuilding replica over its own cells
// [c0, c0 + ncell), starting at its first slot p0 (one workgroup per replica;
// the counts are cleared as they are read), and the end of its last cell,
// cstart[c0 + ncell] = p0 + n (the next replica's first cell start, or
// cstart[C] = N for the last replica).  The replicas that keep their list
// keep their cell starts.
__global__ void __launch_bounds__(DP_SCAN_THREADS) k_scan_reps(i32* count, i32* cstart,
                                                               const DpRep* __restrict__ reps,
                                                               const i32* __restrict__ rlist, const DpCtl* ctl) {
  if (!ctl->rebuild) return;
  __shared__ i32 wsum[DP_SCAN_THREADS / 64];
  __shared__ i32 btotal;
  const int t = threadIdx.x;
  const int lane = t & 63;
  const int w = t >> 6;
  constexpr int NW = DP_SCAN_THREADS / 64;
  constexpr int PER = 4;
  const int nr = ctl->rb_n;
  for (int j = blockIdx.x; j < nr; j += gridDim.x) {
    const DpRep& P = reps[rlist[j]];
    i32* cnt = count + P.c0;
    i32* cst = cstart + P.c0;
    const i64 C = P.ncell;
    i32 carry = (i32)P.p0;
    for (i64 b0 = 0; b0 < C; b0 += (i64)PER * DP_SCAN_THREADS) {
      const i64 c0 = b0 + (i64)PER * t;
      i32 v[PER];
      i32 sthr = 0;
#pragma unroll
      for (int k = 0; k < PER; ++k) {
        v[k] = (c0 + k < C) ? cnt[c0 + k] : 0;
        sthr += v[k];
      }
      i32 incl = sthr;
#pragma unroll
      for (int off = 1; off < 64; off <<= 1) {
        const i32 y = __shfl_up(incl, off);
        if (lane >= off) incl += y;
      }
      if (lane == 63) wsum[w] = incl;
      __syncthreads();
      if (w == 0) {
        const i32 xw = (lane < NW) ? wsum[lane] : 0;
        i32 xi = xw;
#pragma unroll
        for (int off = 1; off < NW; off <<= 1) {
          const i32 y = __shfl_up(xi, off);
          if (lane >= off) xi += y;
        }
        if (lane < NW) wsum[lane] = xi - xw;
        if (lane == NW - 1) btotal = xi;
      }
      __syncthreads();
      i32 ex = carry + wsum[w] + (incl - sthr);
#pragma unroll
      for (int k = 0; k < PER; ++k) {
        if (c0 + k < C) {
          cst[c0 + k] = ex;
          ex += v[k];
          cnt[c0 + k] = 0;
        }
      }
      carry += btotal;
      __syncthreads();  // wsum / btotal are rewritten by the next block
    }
    if (t == 0) cst[C] = (i32)(P.p0 + P.n);
  }
}

// Emit for the tiles marked by k_emit_tile (halo larger than the LDS record
// area): the same emit reading records from global memory.
template <int D, typename E>
__global__ void __launch_bounds__(DP_TILE_THREADS) k_emit_tile_far(const double* __restrict__ xref,
                                                                   const i32* __restrict__ cellpack, double qinv,
                                                                   const DpRep* __restrict__ reps, i64 ntiles, int K,
                                                                   E* nl, i32* nlcnt, const i32* tiletab, int tstride,
                                                                   DpCtl* ctl, int force,
                                                                   const i32* __restrict__ todo, int efmt) {
  if (!force && !ctl->rebuild) return;
  if (ctl->nonfinite) return;
  // a rebuild whose emit marked no tile (k_bin cleared the flag before it);
  // an emit outside a rebuild (force) always looks at every tile
  if (!force && ctl->far_tiles == 0) return;
  __shared__ DpTileSmem<D> S;
  int st_count = 0;
  const i64 ntodo = todo ? (i64)ctl->emit_ntodo : ntiles;
  for (i64 i = blockIdx.x; i < ntodo; i += gridDim.x) {
    const i64 t = todo ? (i64)todo[i] : i;
    const i32* tt = tiletab + t * (i64)tstride;
    if (tt[6] == 0) continue;  // uniform over the workgroup
    dp_tile_load<D>(S, tt, 0);
    const int efmt_t = S.ebo ? efmt : 0;  // the tile's entry format (k_emit_tile's choice)
    const DpRep& P = reps[__builtin_amdgcn_readfirstlane(S.rep)];  // uniform: scalar loads
    // st_off is not part of the stored tables
    constexpr int n3 = DpPow3<D>::value;
    for (int st = threadIdx.x; st < n3; st += blockDim.x) {
      int sd[D];
      dp_stencil_digits<D>(st, P.order, sd);
      int off = 0, stride = 1;
      if (P.order == 0) {
#pragma unroll
        for (int d = 0; d < D; ++d) {
          if (P.nc[d] != 1) off += sd[d] * stride;
          stride *= P.Th[d];
        }
      } else {
#pragma unroll
        for (int d = D - 1; d >= 0; --d) {
          if (P.nc[d] != 1) off += sd[d] * stride;
          stride *= P.Th[d];
        }
      }
      S.st_off[st] = off;
    }
    __syncthreads();
    if (P.collapsed_mask != 0)
      dp_emit_particles<D, E, false, true>(S, P, xref, cellpack, qinv, nullptr, K, nl, nlcnt, ctl, efmt_t, st_count);
    else
      dp_emit_particles<D, E, false, false>(S, P, xref, cellpack, qinv, nullptr, K, nl, nlcnt, ctl, efmt_t,
                                            st_count);
    __syncthreads();
  }
  int mx = st_count;
#pragma unroll
  for (int off = 32; off >= 1; off >>= 1) mx = max(mx, __shfl_xor(mx, off));
  if ((threadIdx.x & 63) == 0 && mx > 0)
    atomicMax(ctl->slot + (blockIdx.x % DP_STAT_SLOTS) * DP_STAT_STRIDE + DP_STAT_COUNT, mx);
}

__global__ void k_clear_stats(DpCtl* ctl) {
  for (int k = threadIdx.x; k < DP_STAT_SLOTS * DP_STAT_STRIDE; k += blockDim.x) ctl->slot[k] = 0;
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

// Neighbour-list export in the reference layout (neighbor_list.py:14-17: row
// = original id, -1 after the last neighbour): one workgroup per tile decodes the staged-halo indices
// of its home particles' rows through the tile tables.
template <int D, typename E>
__global__ void __launch_bounds__(DP_TILE_THREADS) k_nl_export(const E* nl, const i32* nlcnt, const i32* orig,
                                                               const i32* cstart, const DpRep* reps, int R,
                                                               i64 ntiles, int K, int esh, const i32* tiletab,
                                                               int tstride, i32* out) {
  __shared__ DpTileSmem<D> S;
  for (i64 t = blockIdx.x; t < ntiles; t += gridDim.x) {
    dp_tile_setup<D>(S, reps, R, cstart, t, 0x7fffffff);
    const int sh = tiletab[t * (i64)tstride + 7] ? esh : 0;  // the tile's entry format
    const DpRep& P = reps[__builtin_amdgcn_readfirstlane(S.rep)];  // uniform: scalar loads
    for (int k = S.k0 + threadIdx.x; k < S.k1; k += blockDim.x) {
      const int m = dp_home_of<D>(S, P.nhome, k);
      const int uh = S.home_u[m];
      const i64 slot_i = S.hstart[uh] + (k - S.home_p[m]);
      const i64 row = orig[slot_i];
      const int cnt = nlcnt[slot_i] < K ? nlcnt[slot_i] : K;
      for (int kk = 0; kk < K; ++kk) {
        i32 val = -1;
        if (kk < cnt) {
          const int idx = (int)((u32)nl[slot_i * (i64)((K + 7) & ~7) + kk] >> sh);
          const int u = dp_halo_of<D>(S, P.nhalo, idx);
          val = orig[S.hstart[u] + (idx - S.hbase[u])];
        }
        out[row * K + kk] = val;
      }
    }
    __syncthreads();
  }
}

// ---------------------------------------------------------------------------
// Radial distribution function counts (positions.py:58-87): every ordered
// pair (n, m) of one replica, n == m included, at the distance NumPy
// computes: d = x_n - x_m; d -= L * round(d / L) (half to even); r =
// sqrt(d0^2 + d1^2 + ...), products and sums rounded one at a time, summed
// in dimension order.  np.histogram with explicit edges puts r in bin i when
// edges[i] <= r < edges[i+1], the last bin closed on the right; NaN nowhere.
// The kernels bin the squared distance s instead of r: with sqrt correctly
// rounded and monotone, edges[i] <= sqrt(s) <=> s >= thr[i], where thr[i] is
// the smallest double with sqrt(thr[i]) >= edges[i], and sqrt(s) <= edges[-1]
// <=> s <= u_last, the largest double with sqrt(u_last) <= edges[-1] (both
// found on the host).  No device sqrt, no dependence on its rounding.
// ---------------------------------------------------------------------------
#define DP_RDF_MAXBINS 4096

// Bin of s among the thresholds of one chunk of edges (`closed`: the chunk
// holds the last bin); -1 outside.  Equal edges leave empty bins, as in NumPy.
__device__ __forceinline__ int dp_rdf_bin(double s, const double* thr, int ne, int closed, double u_last) {
  if (!(s >= thr[0])) return -1;  // below, or NaN
  if (!(s < thr[ne - 1]) && !(closed && s <= u_last)) return -1;
  int lo = 0, hi = ne - 2;  // largest i <= ne - 2 with thr[i] <= s
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (thr[mid] <= s) lo = mid; else hi = mid - 1;
  }
  return lo;
}

// d - L * rint(d / L).  The quotient is needed only near |d / L| = 1/2 and
// beyond: below 0.5 - 2^-40 (checked with the product by 1 / L, within a few
// ulps of the quotient) it rounds to zero, and d - L * 0 = d.
__device__ __forceinline__ double dp_rdf_wrap(double t, double L, double inv_L) {
  if (fabs(t * inv_L) < 0.5 - 0x1p-40) return t;
  return dp_sub(t, dp_mul(L, rint(t / L)));
}

template <int D>
__device__ __forceinline__ double dp_rdf_sq(const double* a, const double* b, const DpRep& P, const double* inv_L) {
  double s = 0.0;
#pragma unroll
  for (int d = 0; d < D; ++d) {
    const double t = dp_rdf_wrap(dp_sub(a[d], b[d]), P.L[d], inv_L[d]);
    s = (d == 0) ? dp_mul(t, t) : dp_add(s, dp_mul(t, t));
  }
  return s;
}

// Add the workgroup's LDS counts to the global histogram and clear them.
__device__ __forceinline__ void dp_rdf_flush(u32* hl, int nb, unsigned long long* hist) {
  __syncthreads();
  for (int b = threadIdx.x; b < nb; b += blockDim.x)
    if (hl[b]) {
      atomicAdd(hist + b, (unsigned long long)hl[b]);
      hl[b] = 0;
    }
  __syncthreads();
}

// The 32-bit LDS counts are flushed before the pairs a workgroup has binned
// since its last flush could exceed 2^31.
#define DP_RDF_FLUSH_PAIRS (1ull << 31)

// All pairs, 256 x 256 blocks staged in LDS (any box, any size).
template <int D>
__global__ void __launch_bounds__(256) k_rdf_all(const double* x, i64 N, const DpRep* reps, const double* thr_g,
                                                 int ne, int closed, double u_last, unsigned long long* hist) {
  __shared__ double sx[256 * D];
  __shared__ double thr[DP_RDF_MAXBINS + 1];
  __shared__ u32 hl[DP_RDF_MAXBINS];
  const DpRep& P = reps[0];
  double inv_L[D];
#pragma unroll
  for (int d = 0; d < D; ++d) inv_L[d] = 1.0 / P.L[d];
  const int nb = ne - 1;
  for (int b = threadIdx.x; b < ne; b += blockDim.x) thr[b] = thr_g[b];
  for (int b = threadIdx.x; b < nb; b += blockDim.x) hl[b] = 0;
  const i64 nblk = (N + 255) / 256;
  unsigned long long pending = 0;
  for (i64 pb = blockIdx.x; pb < nblk * nblk; pb += gridDim.x) {
    const i64 ib = pb / nblk, jb = pb % nblk;
    if (pending + 65536ull > DP_RDF_FLUSH_PAIRS) {
      dp_rdf_flush(hl, nb, hist);
      pending = 0;
    }
    pending += 65536ull;
    __syncthreads();
    const i64 j = jb * 256 + threadIdx.x;
#pragma unroll
    for (int d = 0; d < D; ++d) sx[threadIdx.x * D + d] = (j < N) ? x[j * D + d] : 0.0;
    __syncthreads();
    const i64 i = ib * 256 + threadIdx.x;
    if (i < N) {
      double xi[D];
#pragma unroll
      for (int d = 0; d < D; ++d) xi[d] = x[i * D + d];
      const int nj = (int)min((i64)256, N - jb * 256);
      for (int jj = 0; jj < nj; ++jj) {
        const int bin = dp_rdf_bin(dp_rdf_sq<D>(xi, sx + jj * D, P, inv_L), thr, ne, closed, u_last);
        if (bin >= 0) atomicAdd(&hl[bin], 1u);
      }
    }
  }
  dp_rdf_flush(hl, nb, hist);
}

// Largest squared displacement since the last rebuild (x - xref, minimum
// image as in the skin check), one value per workgroup; NaN propagates.
template <int D>
__global__ void __launch_bounds__(DP_BLOCK) k_rdf_disp(const double* x, const double* xref, const DpRep* reps, i64 N,
                                                       double* block_max) {
  __shared__ double sh[DP_BLOCK / 64];
  const DpRep& R = reps[0];
  double m = 0.0;
  DP_FOR(i, N) {
    double s = 0.0;
#pragma unroll
    for (int d = 0; d < D; ++d) {
      double t = dp_sub(x[i * D + d], xref[i * D + d]);
      t = dp_minimage_gt(t, R.L[d], R.halfL[d]);
      s = (d == 0) ? dp_mul(t, t) : dp_add(s, dp_mul(t, t));
    }
    if (!(s <= m)) m = s;
  }
#pragma unroll
  for (int off = 32; off >= 1; off >>= 1) {
    const double o = __shfl_xor(m, off);
    if (!(o <= m)) m = o;
  }
  if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = m;
  __syncthreads();
  if (threadIdx.x == 0) {
    for (int w = 1; w < DP_BLOCK / 64; ++w)
      if (!(sh[w] <= m)) m = sh[w];
    block_max[blockIdx.x] = m;
  }
}

// Cell tiles of the last rebuild: valid when every pair closer than the last
// edge lies in neighbouring cells (last edge + twice the largest displacement
// since the rebuild below the cell width, at least 3 cells per dimension).
// One thread per home particle, the 3^D stencil cells of its halo.
template <int D>
__global__ void __launch_bounds__(DP_TILE_THREADS) k_rdf_cells(const double* x, const DpRep* reps, i64 ntiles,
                                                               const i32* tiletab, int tstride,
                                                               const double* thr_g, int ne,
                                                               int closed, double u_last, unsigned long long* hist) {
  __shared__ DpTileSmem<D> S;
  __shared__ double thr[DP_RDF_MAXBINS + 1];
  __shared__ u32 hl[DP_RDF_MAXBINS];
  const int nb = ne - 1;
  for (int b = threadIdx.x; b < ne; b += blockDim.x) thr[b] = thr_g[b];
  for (int b = threadIdx.x; b < nb; b += blockDim.x) hl[b] = 0;
  constexpr int n3 = DpPow3<D>::value;
  unsigned long long pending = 0;
  for (i64 t = blockIdx.x; t < ntiles; t += gridDim.x) {
    dp_tile_load<D>(S, tiletab + t * (i64)tstride, 0);
    const DpRep& P = reps[__builtin_amdgcn_readfirstlane(S.rep)];  // uniform: scalar loads
    // a home particle meets at most every staged record of the halo
    const unsigned long long tile_pairs = (unsigned long long)(S.k1 - S.k0) * (unsigned long long)S.nrec;
    if (pending + tile_pairs > DP_RDF_FLUSH_PAIRS) {
      dp_rdf_flush(hl, nb, hist);
      pending = 0;
    }
    pending += tile_pairs;
    for (int st = threadIdx.x; st < n3; st += blockDim.x) {
      int sd[D];
      dp_stencil_digits<D>(st, P.order, sd);
      int off = 0, stride = 1;
      if (P.order == 0) {
#pragma unroll
        for (int d = 0; d < D; ++d) { off += sd[d] * stride; stride *= P.Th[d]; }
      } else {
#pragma unroll
        for (int d = D - 1; d >= 0; --d) { off += sd[d] * stride; stride *= P.Th[d]; }
      }
      S.st_off[st] = off;
    }
    __syncthreads();
    for (int k = S.k0 + threadIdx.x; k < S.k1; k += blockDim.x) {
      const int m = dp_home_of<D>(S, P.nhome, k);
      const int uh = S.home_u[m];
      const i64 slot_i = S.hstart[uh] + (k - S.home_p[m]);
      double xi[D], inv_L[D];
#pragma unroll
      for (int d = 0; d < D; ++d) {
        xi[d] = x[slot_i * D + d];
        inv_L[d] = 1.0 / P.L[d];
      }
      for (int st = 0; st < n3; ++st) {
        const int u = uh + S.st_off[st];
        const i64 s0 = S.hstart[u];
        const int cnt = S.hbase[u + 1] - S.hbase[u];
        for (int q = 0; q < cnt; ++q) {
          const int bin = dp_rdf_bin(dp_rdf_sq<D>(xi, x + (s0 + q) * D, P, inv_L), thr, ne, closed, u_last);
          if (bin >= 0) atomicAdd(&hl[bin], 1u);
        }
      }
    }
    __syncthreads();
  }
  dp_rdf_flush(hl, nb, hist);
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
