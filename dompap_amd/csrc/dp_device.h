// dp_device.h -- device-side data layout and helpers shared by the hipcc-built
// pipeline kernels (dp_kernels.h) and the hiprtc-built pair kernels
// (dp_pair.hip).  Plain C structs only, so that the host library, hipcc and
// hiprtc all see the same layout.
#pragma once

#ifdef __HIPCC_RTC__
typedef unsigned int u32;
typedef int i32;
typedef unsigned long long u64;
typedef long long i64;
typedef unsigned short u16;
#else
#include <hip/hip_runtime.h>
#include <stdint.h>
typedef uint32_t u32;
typedef int32_t i32;
typedef uint64_t u64;
typedef int64_t i64;
typedef uint16_t u16;
#endif

#define DP_DMAX 4          // largest spatial dimension supported (stencil rows fit one wavefront)
#define DP_MAXTYPES 8      // largest number of particle types (sigma/eps tables)
#define DP_MAXPOT 8        // largest number of pair potentials in one context (one per replica)
#define DP_HALO_MAX 512    // halo cells per tile
#define DP_HOME_MAX 128    // home cells per tile
#define DP_TILE_THREADS 256
// bytes allocated after the neighbour list: a lane's row walk loads the
// next 16-byte group (8 x lanes-per-particle entries on) before it tests
// whether its row goes on
#define DP_NL_SLACK 4096

// Neighbour-entry format efmt = esh | sent << 8: an entry is a staged-halo
// index shifted left by esh (in mixed precision the byte offset of the
// neighbour's record in the pair kernels' LDS record area); with sent, rows
// are padded with the entry of a sentinel record staged after the halo.  A
// tile takes this format when its halo fits the pair kernels' LDS record
// area at the emit (word 7 of its table); other tiles keep plain indices
// padded with 0 (efmt 0).  Largest staged halo (records) that 16-bit entries
// of a format address: entries reach nrec << esh with a sentinel,
// (nrec - 1) << esh without.
__host__ __device__ __forceinline__ int dp_efmt_hlim(int efmt) {
  const int esh = efmt & 31;
  return (efmt >> 8) ? (65535 >> esh) : (65535 >> esh) + 1;
}

// One independent system ("replica").  Particles of replica r occupy slots
// [p0, p0+n) of every per-particle array; its cells occupy [c0, c0+ncell);
// its tiles (groups of home cells handled by one workgroup) [t0, t0+ntiles).
// Cells and tiles are linearised with the stencil's fastest digit fastest:
// dimension 0 fastest when `order` is 0, dimension D-1 fastest when it is 1.
struct DpRep {
  i64 p0, n;            // particle slot range
  i64 c0, ncell;        // global cell range
  i64 t0, ntiles;       // global tile range
  i32 nc[DP_DMAX];      // cells per dimension (>= 3, or 1 = collapsed)
  i32 T[DP_DMAX];       // home cells per tile per dimension
  i32 nt[DP_DMAX];      // tiles per dimension
  i32 Th[DP_DMAX];      // halo box extent per dimension (T + 2, or 1 when collapsed)
  i32 Wq[DP_DMAX];      // cell width in fixed-point units, rint(w * 2^qbits)
  i32 Lq[DP_DMAX];      // box length in fixed-point units (collapsed dimensions)
  double L[DP_DMAX];    // box lengths
  double halfL[DP_DMAX];// L/2 (reference: box_vectors[d] / 2)
  double w[DP_DMAX];    // cell width L/nc
  double inv_w[DP_DMAX];// nc/L
  double nl_thresh;     // s < nl_thresh  <=>  sqrt(s) < cutoff (exact, FP64)
  u64 nlq_lo, nlq_hi;   // fixed-point r^2 (units^2) below: certainly listed; above: certainly not
  float nlq_lo_f, nlq_hi_f;  // FP32 screen of the same: a FP32 sum of squares below lo_f is below nlq_lo, above hi_f above nlq_hi
  i32 nlq16_lo, nlq16_hi;    // 16-bit screen (offsets >> 16): s16 < lo is below nlq_lo, s16 > hi above nlq_hi
  double skin_half_sq;  // (0.5 * skin)^2
  double dt, T_, tau;   // integrator parameters
  double g1_sq, g1_ca, g1_cb;  // G-JF coefficients for mass 1 (integrator.py:17-40 order, on the host)
  u64 seed;             // noise key of the replica (dp_set_seed / dp_set_replica_seeds)
  i32 nhalo;            // prod(Th)
  i32 nhome;            // prod(T)
  i32 collapsed_mask;   // bit d set if nc[d] == 1
  i32 order;            // 0: dimension 0 fastest, 1: dimension D-1 fastest
  i32 pot;              // potential of the replica (index into the context's pair tables)
  i32 nslice;           // tiles per box of home cells: the home particles split into nslice
                        // ranges, one workgroup each (single-cell replicas of many particles)
};

// Device control block (one per context).  The neighbour-list statistics are
// kept in DP_STAT_SLOTS slots, one 128-byte line each, that workgroups update
// by blockIdx % DP_STAT_SLOTS: same-address device atomics from thousands of
// workgroups serialise, spread ones do not.  The host reduces the slots into
// max_count / max_cell / max_halo after each read-back.
#define DP_STAT_SLOTS 64
#define DP_BLK_MAX 128    // steps per step block, at most
#define DP_STAT_STRIDE 32
struct DpCtl {
  u64 step_counter;     // steps taken by the device (RNG counter)
  u64 step_index;       // counter value used by the current step
  i32 rebuild;          // 1 if the current step rebuilds the neighbour list
  i32 force_rebuild;    // host request
  i32 need_rebuild;     // some particle moved more than half the skin since the last build
  i32 overflow;         // some row needed more than K entries
  i32 nonfinite;        // a non-finite position was seen while binning
  i32 cell_overflow;    // a cell held more particles than the entry format allows
  i32 max_count;        // host-side reduction of the slots: largest row length
  i32 max_cell;         //   largest cell occupancy at the last rebuild
  i32 max_halo;         //   largest staged halo (records) at the last rebuild
  i32 blk_done;         // steps the last step block completed
  i32 blk_perstep;      // 1: the block rebuilds (gated) before every step and never stops early
  i32 emit_ntodo;       // tiles in the emit's tile list (batches: the rebuilding replicas' tiles)
  i32 far_tiles;        // the rebuild's emit marked a tile whose halo exceeds its LDS record area
  i32 pblk_n;           // batches: 256-particle blocks of the rebuilding replicas (k_rebuild_list)
  i32 rb_n;             //   and the number of rebuilding replicas
  i32 pad_[3];
  // step blocks: the rebuild pipeline runs once at the block's start; step p
  // of the block runs while no earlier step of the block raised blk_need and
  // blk_base + p < blk_target (blk_perstep = 0); or the pipeline runs before
  // every step and step p runs while blk_base + p < blk_target (blk_perstep =
  // 1).  Which replicas rebuild is kept per replica outside the block (the
  // context's flag array: need[R] raised by the skin trigger, then
  // rebuild[R] for the current step).
  u64 blk_base;         // step counter at the block's start
  u64 blk_target;       // first step index the block must not run
  u64 blk_left;         // steps the host asks for (set before the block)
  i32 blk_need[DP_BLK_MAX];  // step p of the block found a particle beyond half the skin
  i32 slot[DP_STAT_SLOTS * DP_STAT_STRIDE];  // [s * stride + 0 / 1 / 2]: the three maxima
};
static_assert(__builtin_offsetof(DpCtl, blk_need) % 16 == 0, "blk_need is read 16 bytes at a time");
#define DP_STAT_COUNT 0
#define DP_STAT_CELL 1
#define DP_STAT_HALO 2

// Pair-interaction tables (per context; indexed [ti * ntypes + tj]).
struct DpPairTab {
  double sigma[DP_MAXTYPES * DP_MAXTYPES];
  double eps[DP_MAXTYPES * DP_MAXTYPES];
  double cut_thresh[DP_MAXTYPES * DP_MAXTYPES];   // s > t  <=>  sqrt(s)/sigma > r_cut
  double inv_sigma2[DP_MAXTYPES * DP_MAXTYPES];   // 1/sigma^2
  double eps_over_sigma[DP_MAXTYPES * DP_MAXTYPES];
  float guard_lo_q[DP_MAXTYPES * DP_MAXTYPES];    // fixed-point r^2 below: certainly inside
  float guard_hi_q[DP_MAXTYPES * DP_MAXTYPES];    // fixed-point r^2 at or above: certainly outside
  float inv_sigma2_qf[DP_MAXTYPES * DP_MAXTYPES]; // 2^-2q / sigma^2 (FP32 pair path)
  float eps_over_sigma_qf[DP_MAXTYPES * DP_MAXTYPES];  // eps / sigma * 2^-q (FP32 pair path)
  float eps_f[DP_MAXTYPES * DP_MAXTYPES];
  float sigma_f[DP_MAXTYPES * DP_MAXTYPES];
  double r_cut;
  double u_rc;          // u(r_cut) used for the energy shift
  double qscale;        // 2^-qbits
  double qscale2;       // 2^-2qbits
  double qinv;          // 2^qbits
  // mixed mode, potentials whose f(x)/x is a polynomial in 1/x^2 (at most
  // four terms c_k x^-2e_k): per type pair, the coefficients
  // a_k = c_k * eps / sigma * 2^-q * (sigma^2 2^2q)^e_k of the force term
  // per fixed-point unit as a polynomial in 1/s, s the squared fixed-point
  // distance; `fold` = 1 when every a_k is a normal number
  double gk[DP_MAXTYPES * DP_MAXTYPES * 4];
  i32 fold;
};

// Arguments of the hiprtc-built pair / step kernels (passed by value).
struct DpPairArgs {
  // inputs (current buffers)
  const double* x;        // [N*D] FP64 positions, slot order
  const i32* xq;          // [N*RS] mixed records: fixed-point cell-relative coordinates + type
  const void* nl;         // [N*K] neighbour entries (u16 or u32), row major
  const i32* nlcnt;       // [N]
  const i32* type;        // [N]
  const i32* cstart;      // [C+1] first slot of every cell
  const DpRep* reps;      // [R]
  const DpPairTab* tab;
  i64 N;
  i64 ntiles;             // total tiles (grid size)
  i32 K, R;
  i32 cap;                // records that fit in the workgroup's LDS
  i32 tstride;            // ints per tile in tiletab
  const i32* tiletab;     // [ntiles * tstride] tile tables written at the last build
  // observable outputs (nullable)
  double* F;              // [N*D]
  double* E;              // [N]
  double* W;              // [N]
  double* LAP;            // [N]
  // integrator state (fused step kernels)
  double* x_out;          // [N*D]
  i32* xq_out;            // [N*RS]
  double* v;              // [N*D]
  double* beta;           // [N*D]
  i32* img;               // [N*D]
  const double* mass;     // [N]
  const double* xref;     // [N*D]
  const i32* cellpack;    // [N]
  const i32* orig;        // [N]
  const i32* rep;         // [N]
  const double* xi;       // injected normals [N*D] in original order, or null
  u64 seed;
  DpCtl* ctl;
  i32 wrap;               // 1: wrap positions into the box after the update
  i32 blk_pos;            // position of this step in a step block, -1 outside blocks
  i32* rflag;             // [2R]: need[r] (raised by replica r's skin trigger), rebuild[r]
  i32 unit_mass;          // 1: every mass is 1.0 (the integrators skip reading mass)
};

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

// Two standard normals from 128 random bits (Box-Muller on two 53-bit
// uniforms strictly inside (0,1)).  The angle 2 pi u2 goes through sincospi
// (no range reduction: fewer registers than sincos, within an ulp of it).
__device__ __forceinline__ void dp_box_muller(DpU4 b, double& z0, double& z1) {
  u64 a = ((u64)b.x << 32) | b.y;
  u64 c = ((u64)b.z << 32) | b.w;
  double u1 = ((double)(a >> 11) + 0.5) * 1.1102230246251565404e-16;  // 2^-53
  double u2 = ((double)(c >> 11) + 0.5) * 1.1102230246251565404e-16;
  double rad = sqrt(-2.0 * log(u1));
  double s, co;
  sincospi(2.0 * u2, &s, &co);
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

// Standard normal numbers 2k and 2k + 1 of particle `pid` at step `step`
// (one Philox block and one Box-Muller pair: dp_normal(.., 2k) and
// dp_normal(.., 2k + 1) in one call).
__device__ __forceinline__ void dp_normal2(u64 seed, u64 step, u32 pid, int k, double& z0, double& z1) {
  DpU4 c;
  c.x = pid;
  c.y = (u32)k;
  c.z = (u32)step;
  c.w = (u32)(step >> 32);
  dp_box_muller(dp_philox4x32_10(c, (u32)seed, (u32)(seed >> 32)), z0, z1);
}

// positions.py:90-98 wrap of one coordinate: shift = floor(x / L) (0 without
// the division when 0 < x < L, where the rounded quotient is in (0, 1); zeros
// take the division, which keeps the sign of a -0 shift).
__device__ __forceinline__ double dp_wrap_shift(double x, double L) {
  return (x > 0.0 && x < L) ? 0.0 : floor(x / L);
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

// ------------------------------------------------------------------------
// Tile geometry.  A workgroup handles the home cells of one tile; the halo
// box (home cells plus one layer, periodic) is staged in LDS.  Halo cells are
// linearised in the stencil's digit order, so that the cells of a stencil
// row along the fastest digit are consecutive.
// ------------------------------------------------------------------------
template <int D> struct DpPow3 { static constexpr int value = 3 * DpPow3<D - 1>::value; };
template <> struct DpPow3<0> { static constexpr int value = 1; };

// Capacity of the tile tables in LDS: DP_HALO_MAX / DP_HOME_MAX cells, or, in
// the run-time compiled pair kernels, the plan's largest halo / home cell
// counts (DP_TS_HALO / DP_TS_HOME, defined by the host): 3-D tiles have ~48
// halo cells, and the 7 KB of a 512-cell table would cost a workgroup per CU
// when the staged records are large.
#ifndef DP_TS_HALO
#define DP_TS_HALO DP_HALO_MAX
#endif
#ifndef DP_TS_HOME
#define DP_TS_HOME DP_HOME_MAX
#endif
template <int D>
struct DpTileSmem {
  i32 hstart[DP_TS_HALO];      // first slot of each halo cell
  i32 hcoord[DP_TS_HALO];      // halo-box coordinates of each halo cell, 8 bits per dimension
  i32 hbase[DP_TS_HALO + 1];   // staged-record offset of each halo cell (prefix)
  i32 home_u[DP_TS_HOME];      // halo index of each home cell (-1: outside the box)
  i32 home_p[DP_TS_HOME + 1];  // home-particle prefix over home cells
  i32 st_off[DpPow3<D>::value];// halo-index offset of each stencil entry
  i32 nrec;                    // staged records in the halo
  i32 maxc;                    // largest halo-cell occupancy
  i32 nhomep;                  // home particles
  i32 k0, k1;                  // this tile's range of them (a slice; all of them when nslice == 1)
  i32 rep;                     // replica of the tile
  i32 fits;                    // 1 if the halo fits in the LDS record area
  i32 ebo;                     // 1 if the tile's entries are byte offsets (table word 7)
};

// neighbor_list.py:183-201 trigger for one particle: squared displacement
// beyond (skin / 2)^2; NaN also triggers, so that a rebuild reports it.
__device__ __forceinline__ bool dp_skin_trips(double s, double skin_half_sq) {
  return s > skin_half_sq || s != s;
}

// Raise ctl->need_rebuild if any lane of the wavefront tripped: a plain store
// of 1 by one lane, only from wavefronts that tripped (no same-address atomics).
__device__ __forceinline__ void dp_raise_rebuild(bool trip, i32* flag) {
  const u64 b = __ballot(trip);
  if (b != 0ull && (int)(threadIdx.x & 63) == __ffsll((unsigned long long)b) - 1)
    *(volatile i32*)flag = 1;
}

// Replica r's skin trigger: rare (a few particles per rebuild), so every
// lane that trips stores the flag itself.
__device__ __forceinline__ void dp_raise_replica(bool trip, i32* need, int r) {
  if (trip && need) *(volatile i32*)(need + r) = 1;
}

// Whether step `pos` of a step block runs (always outside blocks, pos < 0).
// The earlier steps' flags are read 4 at a time with no early exit, so the
// loads are independent (scalar loads in kernels that do not write ctl).
__device__ __forceinline__ bool dp_blk_go(const DpCtl* ctl, int pos) {
  if (pos < 0) return true;
  if (ctl->blk_base + (u64)pos >= ctl->blk_target) return false;
  if (ctl->blk_perstep) return true;
  const int4* w = (const int4*)ctl->blk_need;
  i32 any = 0;
  for (int q = 0; q < pos; q += 4) {
    const int4 f = w[q >> 2];
    any |= f.x | ((q + 1 < pos) ? f.y : 0) | ((q + 2 < pos) ? f.z : 0) | ((q + 3 < pos) ? f.w : 0);
  }
  return any == 0;
}

// Step index (noise counter) of the current step, and the flag its skin
// trigger raises.
__device__ __forceinline__ u64 dp_step_index(const DpCtl* ctl, int pos) {
  return pos < 0 ? ctl->step_index : ctl->blk_base + (u64)pos;
}
__device__ __forceinline__ i32* dp_need_flag(DpCtl* ctl, int pos) {
  return pos < 0 ? &ctl->need_rebuild : &ctl->blk_need[pos];
}

// Replica owning global tile t.
__device__ __forceinline__ int dp_tile_replica(const DpRep* reps, int R, i64 t) {
  int lo = 0, hi = R - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (reps[mid].t0 <= t) lo = mid; else hi = mid - 1;
  }
  return lo;
}

// Digits of a linear index in the replica's digit order.
template <int D>
__device__ __forceinline__ void dp_unlinear(i64 lin, const i32* ext, int order, int out[D]) {
  if (order == 0) {
#pragma unroll
    for (int d = 0; d < D; ++d) { out[d] = (int)(lin % ext[d]); lin /= ext[d]; }
  } else {
#pragma unroll
    for (int d = D - 1; d >= 0; --d) { out[d] = (int)(lin % ext[d]); lin /= ext[d]; }
  }
}

template <int D>
__device__ __forceinline__ i64 dp_linear(const int* c, const i32* ext, int order) {
  i64 lin = 0, stride = 1;
  if (order == 0) {
#pragma unroll
    for (int d = 0; d < D; ++d) { lin += (i64)c[d] * stride; stride *= ext[d]; }
  } else {
#pragma unroll
    for (int d = D - 1; d >= 0; --d) { lin += (i64)c[d] * stride; stride *= ext[d]; }
  }
  return lin;
}

// Stencil digits s_d in {-1, 0, 1} of stencil index st.
template <int D>
__device__ __forceinline__ void dp_stencil_digits(int st, int order, int s[D]) {
  if (order == 0) {
#pragma unroll
    for (int d = 0; d < D; ++d) { s[d] = st % 3 - 1; st /= 3; }
  } else {
#pragma unroll
    for (int d = D - 1; d >= 0; --d) { s[d] = st % 3 - 1; st /= 3; }
  }
}

// Fill the tile tables for global tile t (every thread of the workgroup calls
// this; it ends with a barrier).  `cap` is the LDS record capacity.
template <int D>
__device__ void dp_tile_setup(DpTileSmem<D>& S, const DpRep* reps, int R, const i32* cstart, i64 t, int cap) {
  const int tid = threadIdx.x;
  const int r = dp_tile_replica(reps, R, t);
  const DpRep& P = reps[r];
  int tc[D];
  const i64 lt = t - P.t0;
  const i64 ct = lt / P.nslice;  // box of home cells; slice lt - ct * nslice of its particles
  const int sl = (int)(lt - ct * P.nslice);
  dp_unlinear<D>(ct, P.nt, P.order, tc);
  const int nh = P.nhalo;
  for (int u = tid; u < nh; u += blockDim.x) {
    int hc[D];
    dp_unlinear<D>(u, P.Th, P.order, hc);
    int cc[D];
#pragma unroll
    for (int d = 0; d < D; ++d) {
      if (P.nc[d] == 1) {
        cc[d] = 0;
      } else {
        int c = tc[d] * P.T[d] + hc[d] - 1;
        c %= P.nc[d];
        if (c < 0) c += P.nc[d];
        cc[d] = c;
      }
    }
    const i64 cell = P.c0 + dp_linear<D>(cc, P.nc, P.order);
    const i32 s0 = cstart[cell];
    S.hstart[u] = s0;
    i32 pk = 0;
#pragma unroll
    for (int d = 0; d < D; ++d) pk |= hc[d] << (8 * d);
    S.hcoord[u] = pk;
    S.hbase[u + 1] = cstart[cell + 1] - s0;
  }
  const int nhm = P.nhome;
  for (int m = tid; m < nhm; m += blockDim.x) {
    int mc[D];
    dp_unlinear<D>(m, P.T, P.order, mc);
    int hc[D];
    bool valid = true;
#pragma unroll
    for (int d = 0; d < D; ++d) {
      const bool coll = P.nc[d] == 1;
      hc[d] = coll ? 0 : mc[d] + 1;
      if (!coll && tc[d] * P.T[d] + mc[d] >= P.nc[d]) valid = false;
    }
    S.home_u[m] = valid ? (int)dp_linear<D>(hc, P.Th, P.order) : -1;
    int cnt = 0;
    if (valid) {
      int cc[D];
#pragma unroll
      for (int d = 0; d < D; ++d) cc[d] = (P.nc[d] == 1) ? 0 : tc[d] * P.T[d] + mc[d];
      const i64 cell = P.c0 + dp_linear<D>(cc, P.nc, P.order);
      cnt = cstart[cell + 1] - cstart[cell];
    }
    S.home_p[m + 1] = cnt;
  }
  constexpr int n3 = DpPow3<D>::value;
  for (int st = tid; st < n3; st += blockDim.x) {
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
  // exclusive prefixes: wavefront 0 over the halo cells, wavefront 1 over the
  // home cells (in-register scans in chunks of 64, carry kept in a register)
  const int lane = tid & 63;
  if (tid < 64) {
    int carry = 0, mx = 0;
    for (int u0 = 0; u0 < nh; u0 += 64) {
      const int u = u0 + lane;
      const int v = (u < nh) ? S.hbase[u + 1] : 0;
      int incl = v;
#pragma unroll
      for (int off = 1; off < 64; off <<= 1) {
        const int t = __shfl_up(incl, off);
        if (lane >= off) incl += t;
      }
      if (u < nh) S.hbase[u + 1] = carry + incl;
      mx = max(mx, v);
      carry += __shfl(incl, 63);
    }
#pragma unroll
    for (int off = 32; off >= 1; off >>= 1) mx = max(mx, __shfl_xor(mx, off));
    if (lane == 0) {
      S.hbase[0] = 0;
      S.nrec = carry;
      S.maxc = mx;
      S.rep = r;
      S.fits = carry <= cap ? 1 : 0;
    }
  } else if (tid < 128) {
    int carry = 0;
    for (int m0 = 0; m0 < nhm; m0 += 64) {
      const int m = m0 + lane;
      int incl = (m < nhm) ? S.home_p[m + 1] : 0;
#pragma unroll
      for (int off = 1; off < 64; off <<= 1) {
        const int t = __shfl_up(incl, off);
        if (lane >= off) incl += t;
      }
      if (m < nhm) S.home_p[m + 1] = carry + incl;
      carry += __shfl(incl, 63);
    }
    if (lane == 0) {
      S.home_p[0] = 0;
      S.nhomep = carry;
      const int h = (carry + P.nslice - 1) / P.nslice;
      S.k0 = min(sl * h, carry);
      S.k1 = min(S.k0 + h, carry);
    }
  }
  __syncthreads();
}

// Tile tables in global memory, written by the neighbour-list emit at every
// build (cell contents are fixed between builds) and loaded by the per-step
// kernels with one coalesced copy instead of recomputing them:
//   [0] nrec [1] nhomep [2] rep [3] maxc [4] nhalo [5] nhome
//   [6] 1 if the halo exceeds the emit's LDS record area [7] 1 if the
//   tile's neighbour entries are byte offsets of staged records (else plain
//   staged-halo indices),
//   [8] [9] the tile's slice [k0, k1) of the home particles [10] [11] unused,
//   then hstart[nh], hbase[nh + 1], hcoord[nh], home_u[nhm], home_p[nhm + 1].
#define DP_TT_HEAD 12
__host__ __device__ __forceinline__ int dp_tile_table_ints(int nh, int nhm) {
  return DP_TT_HEAD + 3 * nh + 1 + 2 * nhm + 1;
}

template <int D>
__device__ void dp_tile_store(const DpTileSmem<D>& S, const DpRep& P, i32* tt) {
  const int nh = P.nhalo, nhm = P.nhome;
  const int tid = threadIdx.x;
  if (tid == 0) {
    tt[0] = S.nrec;
    tt[1] = S.nhomep;
    tt[2] = S.rep;
    tt[3] = S.maxc;
    tt[4] = nh;
    tt[5] = nhm;
    tt[8] = S.k0;
    tt[9] = S.k1;
  }
  i32* p = tt + DP_TT_HEAD;
  for (int u = tid; u < nh; u += blockDim.x) {
    p[u] = S.hstart[u];
    p[2 * nh + 1 + u] = S.hcoord[u];
  }
  for (int u = tid; u <= nh; u += blockDim.x) p[nh + u] = S.hbase[u];
  for (int m = tid; m < nhm; m += blockDim.x) p[3 * nh + 1 + m] = S.home_u[m];
  for (int m = tid; m <= nhm; m += blockDim.x) p[3 * nh + 1 + nhm + m] = S.home_p[m];
}

// Load tile t's tables (every thread calls this; it ends with a barrier).
// S.fits: the halo fits cap records, and with need_bo the tile's entries are
// byte offsets (the pair kernels' LDS path reads records by them).
template <int D>
__device__ void dp_tile_load(DpTileSmem<D>& S, const i32* tt, int cap, bool need_bo = false) {
  const int tid = threadIdx.x;
  const int4 h0 = *(const int4*)tt;
  const int4 h1 = *(const int4*)(tt + 4);
  const int4 h2 = *(const int4*)(tt + 8);
  const int nh = h1.x, nhm = h1.y;
  const i32* p = tt + DP_TT_HEAD;
  for (int u = tid; u < nh; u += blockDim.x) {
    S.hstart[u] = p[u];
    S.hcoord[u] = p[2 * nh + 1 + u];
  }
  for (int u = tid; u <= nh; u += blockDim.x) S.hbase[u] = p[nh + u];
  for (int m = tid; m < nhm; m += blockDim.x) S.home_u[m] = p[3 * nh + 1 + m];
  for (int m = tid; m <= nhm; m += blockDim.x) S.home_p[m] = p[3 * nh + 1 + nhm + m];
  if (tid == 0) {
    S.nrec = h0.x;
    S.nhomep = h0.y;
    S.k0 = h2.x;
    S.k1 = h2.y;
    S.rep = h0.z;
    S.maxc = h0.w;
    S.ebo = h1.w;
    S.fits = (h0.x <= cap && (!need_bo || h1.w != 0)) ? 1 : 0;
  }
  __syncthreads();
}

// Fixed-point offset of halo cell u's centre from the tile's first home cell
// centre along d: (halo coordinate - 1) * Wq (0 along a collapsed dimension,
// where the single cell's centre is the common origin).  Modulo 2^32: only
// differences of shifted coordinates are used, and those are exact.
template <int D>
__device__ __forceinline__ u32 dp_halo_shift(const DpTileSmem<D>& S, const DpRep& P, int u, int d) {
  if (P.nc[d] == 1) return 0u;
  return (u32)(((S.hcoord[u] >> (8 * d)) & 255) - 1) * (u32)P.Wq[d];
}

// Halo cell holding staged record idx (binary search over hbase).
template <int D>
__device__ __forceinline__ int dp_halo_of(const DpTileSmem<D>& S, int nh, int idx) {
  int lo = 0, hi = nh - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (S.hbase[mid] <= idx) lo = mid; else hi = mid - 1;
  }
  return lo;
}

// Home cell holding home particle k (binary search over home_p).
template <int D>
__device__ __forceinline__ int dp_home_of(const DpTileSmem<D>& S, int nhm, int k) {
  int lo = 0, hi = nhm - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (S.home_p[mid] <= k) lo = mid; else hi = mid - 1;
  }
  return lo;
}

#endif  // DP_NO_DEVICE_HELPERS
