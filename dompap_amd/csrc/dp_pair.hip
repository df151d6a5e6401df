// dp_pair.hip -- pair force / energy / virial / Laplacian kernels and the
// fused force + integrate step kernels.  Compiled at run time with hiprtc for
// gfx950, once per (dimension, precision mode, number of types, potential,
// tile shape); the host prepends
//   #define DP_D <dim>          spatial dimension (1..4)
//   #define DP_MODE <0|1>       0 = FP64 parity mode, 1 = mixed precision
//   #define DP_NTYPES <n>       number of particle types
//   #define DP_NPOT <n>         number of potentials (replicas choose one each)
//   #define DP_RS <n>           ints per mixed-precision record
//   #define DP_TPP <n>          lanes per particle (1, 2, 4, 8, 16)
//   #define DP_NLW <16|32>      neighbour-entry width in bits
//   #define DP_OB <n>           offset bits of an entry (stencil index above)
//   #define DP_PAIRPREC <32|64> mixed mode: precision of the in-range pair terms
//   #define DP_CMASK <mask>     dimensions with a single cell (minimum image in
//                               the pair loop); 0 in the usual case
//   #define DP_ESH <n>          entry = staged index << DP_ESH (mixed mode: the
//                               byte offset of the staged record; 0 in FP64 mode)
//   #define DP_SENT <0|1>       mixed mode: rows are padded with the entry of a
//                               sentinel record staged after the halo
// and, per potential k, namespace dp_p<k> holding DP_EVEN (1 if f(x)/x and
// u(x) are functions of x^2) and the SymPy-generated device functions
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
#ifndef DP_CMASK
#define DP_CMASK 0
#endif
#ifndef DP_ESH
#define DP_ESH 0
#endif
#ifndef DP_SENT
#define DP_SENT 0
#endif

#define DP_N3 (DpPow3<DP_D>::value)
#define DP_OMASK ((1u << DP_OB) - 1u)
#define DP_PACK_BITS ((DP_D == 1) ? 31 : 32 / DP_D)
#define DP_GROUPS (DP_TILE_THREADS / DP_TPP)
#define DP_JMAX ((DP_D + DP_TPP - 1) / DP_TPP)  // dimensions per lane in the integrator

#if DP_NLW == 16
typedef u16 dp_ent_t;
#else
typedef u32 dp_ent_t;
#endif

__device__ __forceinline__ int dp_tt(int ti, int tj) { return ti * DP_NTYPES + tj; }

// Per type-pair constants of the inner loops, hoisted out of them: held in
// registers for one type, in an LDS table for several.
struct DpPC {
  double cut;   // s > cut  <=>  x > r_cut (FP64 s)
  double sig, eps, eos;  // sigma, eps, eps / sigma
  double ks, ke;         // mixed FP64 terms: 2^-2q / sigma^2, eps / sigma * 2^-q
  float ghi, glo;        // mixed screening guards in fixed-point units^2
  float ksf, kef;        // FP32 terms: 2^-2q / sigma^2, eps / sigma * 2^-q
  float epsf, sigf;
  double gk[4];          // mixed FP64 terms, folded form: coefficients of 1/s^e_k (DpPairTab::gk)
};

template <class T>
__device__ __forceinline__ T dp_group_sum(T v) {
#pragma unroll
  for (int off = DP_TPP / 2; off >= 1; off >>= 1) v += __shfl_xor(v, off);
  return v;
}

// The same over a group of tpp lanes (a power of two, the same in the whole
// wavefront): DP_TPP lanes, or more for a tile's last few particles.
template <class T>
__device__ __forceinline__ T dp_group_sum(T v, int tpp) {
  if (tpp == DP_TPP) return dp_group_sum(v);
  for (int off = tpp >> 1; off >= 1; off >>= 1) v += __shfl_xor(v, off);
  return v;
}

// Neighbour rows: stride K rounded up to 8 entries, so that every lane reads
// its entries 8 bytes at a time (4 16-bit or 2 32-bit entries).
__device__ __forceinline__ i64 dp_row_stride(int K) { return (i64)((K + 7) & ~7); }
#define DP_EPL (8 / (int)sizeof(dp_ent_t))  // entries per 8-byte load

// Row walking: each lane loads DP_CPL chunks of DP_EPL entries (8 bytes per
// chunk) at a time, one such load ahead of the chunks being evaluated, so
// that the latency of the row's loads (the row is streamed from HBM once per
// step) overlaps the pair arithmetic of the chunks before them.
#ifndef DP_CPL
#define DP_CPL 2
#endif
struct DpLoad {
  u32 w[2 * DP_CPL];
};
__device__ __forceinline__ DpLoad dp_ld_entries(const dp_ent_t* p) {
  DpLoad v;
#if DP_CPL == 1
  const uint2 t = *(const uint2*)p;
  v.w[0] = t.x; v.w[1] = t.y;
#else
  const uint4 t = *(const uint4*)p;
  v.w[0] = t.x; v.w[1] = t.y; v.w[2] = t.z; v.w[3] = t.w;
#endif
  return v;
}

// The chunks of one load at row entry kj.
template <class Fn>
__device__ __forceinline__ void dp_load_chunks(const DpLoad& cur, int kj, int cnt, Fn& fn) {
#pragma unroll
  for (int j = 0; j < DP_CPL; ++j, kj += DP_EPL) {
    if (j > 0 && kj >= cnt) break;
    int idx[DP_EPL];
#if DP_NLW == 16
    idx[0] = (int)(cur.w[2 * j] & 0xffffu);
    idx[1] = (int)(cur.w[2 * j] >> 16);
    idx[2] = (int)(cur.w[2 * j + 1] & 0xffffu);
    idx[3] = (int)(cur.w[2 * j + 1] >> 16);
#else
    idx[0] = (int)cur.w[2 * j];
    idx[1] = (int)cur.w[2 * j + 1];
#endif
    fn(idx, min(DP_EPL, cnt - kj));
  }
}

// Call fn(idx, nvalid) for the chunks of a row of length cnt that belong to
// lane `sub` of its group of tpp lanes: loads of DP_CPL chunks at entries
// [W * (sub + j * tpp), + W), W = DP_EPL * DP_CPL, for j = 0, 1, ...; idx[c]
// for c >= nvalid is 0: the emit pads every row with zero entries up to a
// multiple of DP_EPL, and a row's last chunk ends within that padding.  The
// next load is issued before the current chunks are evaluated, whether or not
// the row goes on (it then reads the next rows or the list's tail slack,
// DP_NL_SLACK).  With PP the two loads alternate between two register sets
// (no copies, two instances of fn's code: for the hot loop only; in the other
// instances the code would outgrow what the compiler inlines into the kernel).
template <bool PP, class Fn>
__device__ __forceinline__ void dp_for_chunks(const dp_ent_t* row, int cnt, int sub, int tpp, Fn& fn) {
  constexpr int W = DP_EPL * DP_CPL;
  const int step = W * tpp;
  int k0 = sub * W;
  if (k0 >= cnt) return;
  DpLoad a = dp_ld_entries(row + k0);
  for (;;) {
    const DpLoad b = dp_ld_entries(row + k0 + step);
    dp_load_chunks(a, k0, cnt, fn);
    k0 += step;
    if (k0 >= cnt) break;
    if (PP) {
      a = dp_ld_entries(row + k0 + step);
      dp_load_chunks(b, k0, cnt, fn);
      k0 += step;
      if (k0 >= cnt) break;
    } else {
      a = b;
    }
  }
}

// integrator.py:17-40 (G-JF) coefficients of mass m, numpy's left-to-right
// evaluation order; a unit mass takes the host's values (same operations).
__device__ __forceinline__ void dp_gjf_coeffs(const DpRep& P, double m, double& sq, double& ca, double& cb) {
#pragma clang fp contract(off)
  if (m == 1.0) {
    sq = P.g1_sq;
    ca = P.g1_ca;
    cb = P.g1_cb;
  } else {
    const double dt = P.dt;
    const double alpha = m / P.tau;
    const double bvar = 2.0 * alpha * P.T_ * dt;
    sq = sqrt(bvar);
    const double qq = alpha * dt / 2.0 / m;
    const double num = 1.0 - qq;
    const double den = 1.0 + qq;
    ca = num / den;
    cb = 1.0 / den;
  }
}

// One dimension of the G-JF velocity update,
//   v1 = ca * v0 + cb * dt * f / m + cb * (b0 + nb) / 2 / m,
// evaluated left to right; dividing by a unit mass is exact, so m == 1 skips it.
__device__ __forceinline__ double dp_gjf_v(double ca, double cb, double dt, double m, double v0, double f,
                                           double b0, double nb) {
#pragma clang fp contract(off)
  double t1 = cb * dt * f;
  double t2 = cb * (b0 + nb) / 2.0;
  if (m != 1.0) {
    t1 = t1 / m;
    t2 = t2 / m;
  }
  return ca * v0 + t1 + t2;
}

// integrator.py:5-14 (leap-frog) velocity update v + dt * f / m.
__device__ __forceinline__ double dp_lf_v(double dt, double m, double v0, double f) {
#pragma clang fp contract(off)
  const double t = dt * f;
  return v0 + ((m != 1.0) ? t / m : t);
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

// One mixed-precision record: DP_RS ints (fixed-point coordinates, type,
// padding), loaded and stored as one vector.
struct DpRec {
  i32 v[DP_RS];
};
__device__ __forceinline__ DpRec dp_ld_rec(const i32* p) {
  DpRec r;
#if DP_RS == 2
  const int2 t = *(const int2*)p;
  r.v[0] = t.x; r.v[1] = t.y;
#elif DP_RS == 4
  const int4 t = *(const int4*)p;
  r.v[0] = t.x; r.v[1] = t.y; r.v[2] = t.z; r.v[3] = t.w;
#else
  const int4 t0 = *(const int4*)p;
  const int4 t1 = *(const int4*)(p + 4);
  r.v[0] = t0.x; r.v[1] = t0.y; r.v[2] = t0.z; r.v[3] = t0.w;
  r.v[4] = t1.x; r.v[5] = t1.y; r.v[6] = t1.z; r.v[7] = t1.w;
#endif
  return r;
}
__device__ __forceinline__ void dp_st_rec(i32* p, const DpRec& r) {
#if DP_RS == 2
  *(int2*)p = make_int2(r.v[0], r.v[1]);
#elif DP_RS == 4
  *(int4*)p = make_int4(r.v[0], r.v[1], r.v[2], r.v[3]);
#else
  *(int4*)p = make_int4(r.v[0], r.v[1], r.v[2], r.v[3]);
  *(int4*)(p + 4) = make_int4(r.v[4], r.v[5], r.v[6], r.v[7]);
#endif
}

// Staged (LDS) records: the coordinates, and the type only with several
// types, in DP_LS ints (a power of two, at most DP_RS).  With one type this
// halves the record for D = 2 and D = 4 (the HBM record keeps its type slot).
#define DP_LS ((DP_D + (DP_NTYPES > 1 ? 1 : 0)) <= 1 ? 1 : (DP_D + (DP_NTYPES > 1 ? 1 : 0)) <= 2 ? 2 : (DP_D + (DP_NTYPES > 1 ? 1 : 0)) <= 4 ? 4 : 8)
static_assert(DP_LS <= DP_RS, "staged record larger than the HBM record");
#if DP_MODE == 1
static_assert(4 * DP_LS == (1 << DP_ESH), "mixed-mode entries are byte offsets of staged records");
#else
static_assert(DP_ESH == 0 && DP_SENT == 0, "FP64-mode entries are staged indices, padded with 0");
#endif
// First DP_LS ints of a record (the remaining slots read as 0).
__device__ __forceinline__ DpRec dp_ld_lrec(const i32* p) {
  DpRec r;
#pragma unroll
  for (int q = 0; q < DP_RS; ++q) r.v[q] = 0;
#if DP_LS == 1
  r.v[0] = p[0];
#elif DP_LS == 2
  const int2 t = *(const int2*)p;
  r.v[0] = t.x; r.v[1] = t.y;
#elif DP_LS == 4
  const int4 t = *(const int4*)p;
  r.v[0] = t.x; r.v[1] = t.y; r.v[2] = t.z; r.v[3] = t.w;
#else
  r = dp_ld_rec(p);
#endif
  return r;
}
__device__ __forceinline__ void dp_st_lrec(i32* p, const DpRec& r) {
#if DP_LS == 1
  p[0] = r.v[0];
#elif DP_LS == 2
  *(int2*)p = make_int2(r.v[0], r.v[1]);
#elif DP_LS == 4
  *(int4*)p = make_int4(r.v[0], r.v[1], r.v[2], r.v[3]);
#else
  dp_st_rec(p, r);
#endif
}

// Record of the particle in halo cell u at slot `slot`, with its coordinates
// moved from its cell's centre to the tile's first home cell centre.
__device__ __forceinline__ DpRec dp_halo_rec(const DpPairArgs& a, const DpTileSmem<DP_D>& S, const DpRep& P,
                                             int u, i64 slot) {
  DpRec r = dp_ld_rec(a.xq + slot * DP_RS);
#pragma unroll
  for (int d = 0; d < DP_D; ++d) r.v[d] = (i32)((u32)r.v[d] + dp_halo_shift<DP_D>(S, P, u, d));
  return r;
}

// Slot of staged halo record idx.
__device__ __forceinline__ i64 dp_slot_of(const DpTileSmem<DP_D>& S, int nh, int idx) {
  const int u = dp_halo_of<DP_D>(S, nh, idx);
  return (i64)S.hstart[u] + (idx - S.hbase[u]);
}

// The DP_D doubles of one particle, two at a time (8-byte aligned pairs:
// 16-byte loads).
typedef double dp_d2a8 __attribute__((ext_vector_type(2), aligned(8)));
__device__ __forceinline__ void dp_ld_vec(const double* p, double (&v)[DP_D]) {
#pragma unroll
  for (int d = 0; d + 1 < DP_D; d += 2) {
    const dp_d2a8 t = *(const dp_d2a8*)(p + d);
    v[d] = t.x; v[d + 1] = t.y;
  }
  if (DP_D & 1) v[DP_D - 1] = p[DP_D - 1];
}

// Halo records staged per thread and batch (mixed mode; one for D = 4, where
// a batch would raise the kernel's register count past 128).
#ifndef DP_SB
#define DP_SB (DP_D <= 3 ? 3 : 1)
#endif

// ===========================================================================
// Potential-dependent part.  The host compiles it once per potential of the
// context, each copy inside its own namespace dp_p<k> together with that
// potential's generated device functions and DP_EVEN; the tile kernels below
// call the copy of the tile's replica's potential (DP_POT_CALL).
// ===========================================================================
//@@POTENTIAL_BEGIN
#ifndef DP_GS_SCALE
#define DP_GS_SCALE 1.0  // no scaled form generated (odd potentials)
#endif
// Mixed mode evaluates the force term as DP_GS_SCALE * dp_pot_gsn(s): the
// constant goes into ke / kef here.
#if DP_EVEN
#define DP_GSN(T, s) dp_pot_gsn<T>(s)
#define DP_GSK DP_GS_SCALE
#else
#define DP_GSK 1.0
#endif

__device__ __forceinline__ DpPC dp_pc(const DpPairTab& T, int tt) {
  DpPC c;
  c.cut = T.cut_thresh[tt];
  c.sig = T.sigma[tt];
  c.eps = T.eps[tt];
  c.eos = T.eps_over_sigma[tt];
  c.ks = T.qscale2 * T.inv_sigma2[tt];
  c.ke = T.eps_over_sigma[tt] * T.qscale * DP_GSK;
  c.ghi = T.guard_hi_q[tt];
  c.glo = T.guard_lo_q[tt];
  c.ksf = T.inv_sigma2_qf[tt];
  c.kef = T.eps_over_sigma_qf[tt] * (float)DP_GSK;
  c.epsf = T.eps_f[tt];
  c.sigf = T.sigma_f[tt];
#if DP_FOLD
#pragma unroll
  for (int k = 0; k < DP_GSK_N; ++k) c.gk[k] = T.gk[tt * 4 + k];
#endif
  return c;
}

#if DP_NTYPES == 1
#define DP_PC_DECL const DpPC pc0 = dp_pc(T, 0);
#define DP_PC(tt) pc0
#else
#define DP_PC_DECL
#define DP_PC(tt) spc[tt]
#endif



// Per-particle pass over the home particles of the staged tile.  LDS selects
// where neighbour data come from: the staged halo (LDS = true) or global
// memory through the tile tables (tiles whose halo exceeds the LDS capacity).
// KIND: 0 forces, 1 observables (forces + per-particle energy / virial /
// Laplacian), 2 fused force + G-JF, 3 fused force + leap-frog.
template <int KIND, bool LDS, int CM>
__device__ __forceinline__ void dp_tile_particles(const DpPairArgs& a, const DpTileSmem<DP_D>& S, const DpRep& P,
                                                  const void* stage, const DpPC* spc, bool& trip) {
  const DpPairTab& T = a.tab[P.pot];  // this replica's potential
  (void)spc;
  DP_PC_DECL
  const int tid = threadIdx.x;
  const dp_ent_t* nl = (const dp_ent_t*)a.nl;
  const int nh = P.nhalo;
#if DP_MODE == 1
  const i32* srec = (const i32*)stage;
#else
  const double* sxd = (const double*)stage;
  const i32* stype = (const i32*)(sxd + (size_t)a.cap * DP_D);
#endif
  for (int base = S.k0; base < S.k1; base += DP_GROUPS) {
    // lanes: DP_TPP per particle; in a pass of fewer than DP_GROUPS
    // particles, the particles beyond the wavefronts that it fills take the
    // remaining wavefronts with tpp lanes each (a power of two, >= DP_D, the
    // same in the whole wavefront), so that a wavefront holding a few
    // particles does not walk their rows DP_TPP lanes at a time
    int tpp = DP_TPP, group = tid / DP_TPP, sub = tid % DP_TPP;
    const int np = min(S.k1 - base, DP_GROUPS);
    if (np < DP_GROUPS) {
      constexpr int PW = 64 / DP_TPP;  // particles per wavefront at DP_TPP lanes
      const int wf = np / PW;           // wavefronts filled at DP_TPP lanes
      const int nt = np - wf * PW;      // particles left
      const int nw = DP_TILE_THREADS / 64 - wf;
      const int wave = tid >> 6;
      int t = DP_TPP;
      while (t < 64 && nt * 2 * t <= 64 * nw) t *= 2;
      if (wave >= wf && nt > 0 && t >= DP_D && t > DP_TPP) {
        const int lt = tid - wf * 64;
        tpp = t;
        group = wf * PW + lt / t;
        sub = lt % t;
      }
    }
    const int k = base + group;
    const bool active = group < np;
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
    const dp_ent_t* row = nl + slot_i * dp_row_stride(a.K);
    double f[DP_D];
#pragma unroll
    for (int d = 0; d < DP_D; ++d) f[d] = 0.0;
    double e = 0.0, w = 0.0, lap = 0.0;
#if DP_MODE == 1
    // ---------------- mixed: fixed-point displacements ----------------
    const DpRec ri = LDS ? dp_ld_lrec(srec + self * DP_LS) : dp_halo_rec(a, S, P, uh, slot_i);
    const int ti = (DP_NTYPES > 1) ? ri.v[DP_D] : 0;
    // one chunk of DP_EPL entries at a time, branch-free apart from the rare
    // exact decision near the cut: records first, then independent pair terms
    auto chunk = [&](const int (&idx)[DP_EPL], const int nvalid) {
      DpRec rj[DP_EPL];
#pragma unroll
      for (int c = 0; c < DP_EPL; ++c) {
        // entries past the row's end are the emit's padding (the sentinel's
        // entry, or 0), read like the others and excluded from the sums
        // below: by the sentinel's distance, or by their position
        if (LDS) {
          rj[c] = dp_ld_lrec((const i32*)((const char*)srec + idx[c]));  // entry = byte offset
        } else {
          const int ic = (c < nvalid) ? (S.ebo ? (idx[c] >> DP_ESH) : idx[c]) : 0;
          const int u = dp_halo_of<DP_D>(S, nh, ic);
          rj[c] = dp_halo_rec(a, S, P, u, (i64)S.hstart[u] + (ic - S.hbase[u]));
        }
      }
      double dq[DP_EPL][DP_D];  // exact: 32-bit fixed-point differences
      double sd[DP_EPL];        // r^2 in fixed-point units^2 (relative error ~2^-52)
      int tt[DP_EPL];
      bool in[DP_EPL];
      bool nr[DP_EPL];
      bool near = false;
#pragma unroll
      for (int c = 0; c < DP_EPL; ++c) {
#pragma unroll
        for (int d = 0; d < DP_D; ++d) {
          i32 dd = (i32)((u32)ri.v[d] - (u32)rj[c].v[d]);
          if (CM & (1 << d)) {
            // single cell along d (offsets from the box centre): minimum image
            // (a no-op for other replicas, whose displacements are < L/2)
            const i32 h = P.Lq[d] >> 1;
            dd = (dd > h) ? dd - P.Lq[d] : ((dd < -h) ? dd + P.Lq[d] : dd);
          }
          dq[c][d] = (double)dd;
        }
        sd[c] = dq[c][0] * dq[c][0];
#pragma unroll
        for (int d = 1; d < DP_D; ++d) sd[c] = fma(dq[c][d], dq[c][d], sd[c]);
        tt[c] = dp_tt(ti, (DP_NTYPES > 1) ? rj[c].v[DP_D] : 0);
        const DpPC& pc = DP_PC(tt[c]);
        // the guards bound the FP32-rounded value; the FP64 one lies closer
        // (a padding entry: beyond ghi when it names the sentinel)
        in[c] = (DP_SENT && LDS ? true : (c < nvalid)) && (sd[c] < (double)pc.ghi);  // else certainly beyond the cut
        nr[c] = in[c] && (sd[c] > (double)pc.glo);
        near |= nr[c];
      }
      if (near) {
        // too close to the cut to decide from the fixed-point values: decide
        // with the reference's FP64 operations (rare)
#pragma unroll
        for (int c = 0; c < DP_EPL; ++c) {
          if (!nr[c]) continue;
          const i64 slot_j = dp_slot_of(S, nh, idx[c] >> DP_ESH);
          double s = 0.0;
#pragma unroll
          for (int d = 0; d < DP_D; ++d) {
            const double tq = dp_minimage_gt(dp_sub(a.x[slot_i * DP_D + d], a.x[slot_j * DP_D + d]), P.L[d],
                                             P.halfL[d]);
            s = (d == 0) ? dp_mul(tq, tq) : dp_add(s, dp_mul(tq, tq));
          }
          in[c] = !(s > DP_PC(tt[c]).cut);
        }
      }
#pragma unroll
      for (int c = 0; c < DP_EPL; ++c) {
#if DP_PAIRPREC == 64
        const DpPC& pc = DP_PC(tt[c]);
        // entries off the list (padding: s = 0; beyond the cut) are evaluated
        // too and their term is selected away (no trap on inf / NaN)
#if DP_FOLD
        // eps f(x)/r per fixed-point unit straight from the squared
        // fixed-point distance: the scales are in the coefficients; summed
        // only where the entry is in range (no zero term for the others)
        if (KIND != 1) {
          if (in[c]) {
            const double fs = dp_pot_gsk<double>(sd[c], pc.gk);
#pragma unroll
            for (int d = 0; d < DP_D; ++d) f[d] = fma(fs, dq[c][d], f[d]);
          }
          continue;
        }
        const double fs = in[c] ? dp_pot_gsk<double>(sd[c], pc.gk) : 0.0;
#else
        const double sx = sd[c] * pc.ks;
#if DP_EVEN
        const double gn = DP_GSN(double, sx);  // f(x)/x / DP_GS_SCALE
#else
        const double xr = sqrt(sx);
        const double gn = dp_pot_f<double>(xr) * dp_rcp(xr);
#endif
        const double fs = in[c] ? pc.ke * gn : 0.0;  // eps f(x)/r per fixed-point unit
#endif
#pragma unroll
        for (int d = 0; d < DP_D; ++d) f[d] = fma(fs, dq[c][d], f[d]);
        if (KIND == 1 && in[c]) {
#if DP_FOLD
          const double sx = sd[c] * pc.ks;
          const double gn = DP_GSN(double, sx);
#endif
          const double g = gn * DP_GSK;
#if DP_EVEN
          e += pc.eps * (dp_pot_us<double>(sx) - T.u_rc);
#else
          e += pc.eps * (dp_pot_u<double>(xr) - T.u_rc);
#endif
          w += pc.eps * pc.sig * g * sx;
          lap += pc.eps * dp_pot_c<double>(sqrt(sx)) - (double)(DP_D - 1) * pc.eos * g;
        }
#else
        const DpPC& pc = DP_PC(tt[c]);
        const float sxf = (float)sd[c] * pc.ksf;
#if DP_EVEN
        const float gn = DP_GSN(float, sxf);
#else
        const float xrf = sqrtf(sxf);
        const float gn = dp_pot_f<float>(xrf) * dp_rcp(xrf);
#endif
        // FP32 pair term, accumulated in FP64 with the exact displacement
        const double fs = in[c] ? (double)(pc.kef * gn) : 0.0;
#pragma unroll
        for (int d = 0; d < DP_D; ++d) f[d] = fma(fs, dq[c][d], f[d]);
        if (KIND == 1 && in[c]) {
          const float g = gn * (float)DP_GSK;
#if DP_EVEN
          e += (double)(pc.epsf * dp_pot_us<float>(sxf)) - (double)pc.epsf * T.u_rc;
#else
          e += (double)(pc.epsf * dp_pot_u<float>(xrf)) - (double)pc.epsf * T.u_rc;
#endif
          w += (double)(pc.epsf * pc.sigf * g * sxf);
          lap += (double)(pc.epsf * dp_pot_c<float>(sqrtf(sxf))) -
                 (double)(DP_D - 1) * (double)(pc.epsf / pc.sigf * g);
        }
#endif
      }
    };
    dp_for_chunks<LDS && CM == 0 && KIND == 0>(row, cnt, sub, tpp, chunk);
#else
    // ---------------- FP64 parity mode ----------------
    double xi_[DP_D];
#pragma unroll
    for (int d = 0; d < DP_D; ++d) xi_[d] = LDS ? sxd[self * DP_D + d] : a.x[slot_i * DP_D + d];
    const int ti = (DP_NTYPES > 1) ? (LDS ? stype[self] : a.type[slot_i]) : 0;
    // one chunk of DP_EPL entries at a time: positions first, then the
    // reference's FP64 operations, accumulated in row order within the lane
    auto chunk = [&](const int (&idx)[DP_EPL], const int nvalid) {
      double xj[DP_EPL][DP_D];
      int tj[DP_EPL];
#pragma unroll
      for (int c = 0; c < DP_EPL; ++c) {
        const int ic = idx[c];  // past the row's end: the emit's zero padding
        tj[c] = 0;
        if (LDS) {
#pragma unroll
          for (int d = 0; d < DP_D; ++d) xj[c][d] = sxd[ic * DP_D + d];
          if (DP_NTYPES > 1) tj[c] = stype[ic];
        } else {
          const i64 slot_j = dp_slot_of(S, nh, ic);
#pragma unroll
          for (int d = 0; d < DP_D; ++d) xj[c][d] = a.x[slot_j * DP_D + d];
          if (DP_NTYPES > 1) tj[c] = a.type[slot_j];
        }
      }
#pragma unroll
      for (int c = 0; c < DP_EPL; ++c) {
        double dd[DP_D];
        double s = 0.0;
#pragma unroll
        for (int d = 0; d < DP_D; ++d) {
          const double tq = dp_minimage_gt(dp_sub(xi_[d], xj[c][d]), P.L[d], P.halfL[d]);
          dd[d] = tq;
          s = (d == 0) ? dp_mul(tq, tq) : dp_add(s, dp_mul(tq, tq));
        }
        const DpPC& pc = DP_PC(dp_tt(ti, tj[c]));
        // x > r_cut  <=>  s > threshold; off-list entries evaluate at s = 1
        const bool in = (c < nvalid) && !(s > pc.cut);
        const double rr = sqrt(in ? s : 1.0);
        const double xr = rr / pc.sig;
        const double eps = pc.eps;
        const double sf = eps * dp_pot_f<double>(xr);
#pragma unroll
        for (int d = 0; d < DP_D; ++d) f[d] = in ? f[d] + sf * (dd[d] / rr) : f[d];
        if (KIND == 1 && in) {
          e += eps * (dp_pot_u<double>(xr) - T.u_rc);
          w += sf * rr;
          lap += eps * dp_pot_c<double>(xr) + (double)(DP_D - 1) * (-sf) / rr;
        }
      }
    };
    dp_for_chunks<LDS && CM == 0 && KIND == 0>(row, cnt, sub, tpp, chunk);
#endif
    // combine the partial sums of the particle's DP_TPP lanes (every lane
    // of the group ends with the totals)
#pragma unroll
    for (int d = 0; d < DP_D; ++d) f[d] = dp_group_sum(f[d], tpp);
    if (KIND == 1) {
      e = dp_group_sum(e, tpp);
      w = dp_group_sum(w, tpp);
      lap = dp_group_sum(lap, tpp);
    }
    if (!active) continue;
    if (KIND == 0 || KIND == 1) {
      if (a.F) {
#pragma unroll
        for (int d = 0; d < DP_D; ++d)
          if (d % tpp == sub) a.F[slot_i * DP_D + d] = f[d];
      }
      if (KIND == 1 && sub == 0) {
        if (a.E) a.E[slot_i] = e;
        if (a.W) a.W[slot_i] = w;
        if (a.LAP) a.LAP[slot_i] = lap;
      }
      continue;
    }
    // fused integrator update of particle slot_i: lane `sub` of the group
    // updates dimensions sub, sub + tpp, ...; the normal numbers are drawn
    // in one SIMD pass when tpp >= DP_D.  A lane's j-th dimension is held at
    // index j: d / DP_TPP, or 0 where tpp > DP_TPP (then tpp >= DP_D: one
    // dimension per lane)
    {
#pragma clang fp contract(off)
      const i64 ii = slot_i;
      const double m = a.unit_mass ? 1.0 : a.mass[ii];
      const double dt = P.dt;
      double z_own[DP_JMAX];
      if (KIND == 2) {
        const u32 pid = (u32)a.orig[ii];
#pragma unroll
        for (int j = 0; j < DP_JMAX; ++j) {
          const int d = sub + j * tpp;
          z_own[j] = 0.0;
          // device noise: keyed by the replica's seed and the particle's id in its replica
          if (d < DP_D)
            z_own[j] = a.xi ? a.xi[(i64)pid * DP_D + d]
                            : dp_normal(P.seed, dp_step_index(a.ctl, a.blk_pos), (u32)(pid - (u32)P.p0), d);
        }
      }
      double sq = 0.0, ca = 0.0, cb = 0.0;
      if (KIND == 2) dp_gjf_coeffs(P, m, sq, ca, cb);
#if DP_MODE == 1
      const u32 pack = (u32)a.cellpack[ii];
#endif
      double tq_own[DP_JMAX];
#pragma unroll
      for (int j = 0; j < DP_JMAX; ++j) tq_own[j] = 0.0;
#pragma unroll
      for (int d = 0; d < DP_D; ++d) {
        if (d % tpp != sub) continue;
        const int jd = (tpp == DP_TPP) ? d / DP_TPP : 0;
        double xn;
        if (KIND == 2) {
          const double nb = sq * z_own[jd];
          const double v1 = dp_gjf_v(ca, cb, dt, m, a.v[ii * DP_D + d], f[d], a.beta[ii * DP_D + d], nb);
          xn = a.x[ii * DP_D + d] + v1 * dt;
          a.v[ii * DP_D + d] = v1;
          a.beta[ii * DP_D + d] = nb;
        } else {
          // integrator.py:5-14 (leap-frog)
          const double v1 = dp_lf_v(dt, m, a.v[ii * DP_D + d], f[d]);
          xn = a.x[ii * DP_D + d] + dt * v1;
          a.v[ii * DP_D + d] = v1;
        }
        // positions.py:90-98 wrap, then neighbor_list.py:183-201 displacement
        if (a.wrap) {
          const double sh = dp_wrap_shift(xn, P.L[d]);
          if (!(xn > 0.0 && xn < P.L[d])) {
            a.img[ii * DP_D + d] += (i32)sh;
            xn = xn - sh * P.L[d];
          }
        }
        a.x_out[ii * DP_D + d] = xn;
        tq_own[jd] = dp_minimage_gt(dp_sub(xn, a.xref[ii * DP_D + d]), P.L[d], P.halfL[d]);
#if DP_MODE == 1
        const u32 c = (pack >> (DP_PACK_BITS * d)) & ((1u << DP_PACK_BITS) - 1u);
        const double cen = ((double)c + 0.5) * P.w[d];
        a.xq_out[ii * DP_RS + d] = (i32)rint(dp_minimage_gt(xn - cen, P.L[d], P.halfL[d]) * T.qinv);
#endif
      }
#if DP_MODE == 1
      if (sub == 0) {
        // slots after the coordinates: the type (0 with one type), padding
#pragma unroll
        for (int q = DP_D; q < DP_RS; ++q) a.xq_out[ii * DP_RS + q] = (DP_NTYPES > 1) ? a.xq[ii * DP_RS + q] : 0;
      }
#endif
      // squared displacement summed in dimension order from the owners
      const int gbase = (tid & 63) & ~(tpp - 1);
      double sdisp = 0.0;
#pragma unroll
      for (int d = 0; d < DP_D; ++d) {
        const int jd = (tpp == DP_TPP) ? d / DP_TPP : 0;
        const double tq = (tpp == 1) ? tq_own[d] : __shfl(tq_own[jd], gbase + (d % tpp));
        sdisp = (d == 0) ? dp_mul(tq, tq) : dp_add(sdisp, dp_mul(tq, tq));
      }
      trip |= dp_skin_trips(sdisp, P.skin_half_sq);
    }
  }
}

// This potential's type-pair constants into the workgroup's LDS table.
__device__ __forceinline__ void dp_fill_spc(DpPC* spc, const DpPairTab& T) {
  if ((int)threadIdx.x < DP_NTYPES * DP_NTYPES) spc[threadIdx.x] = dp_pc(T, (int)threadIdx.x);
}
#undef DP_GSN
#undef DP_GSK
#undef DP_GS_SCALE
#undef DP_PC_DECL
#undef DP_PC
//@@POTENTIAL_END
// ===========================================================================

template <int KIND>
__device__ void dp_tile_run(const DpPairArgs& a) {
  if (!dp_blk_go(a.ctl, a.blk_pos)) return;  // a skipped step of a step block (uniform)
  __shared__ DpTileSmem<DP_D> S;
  extern __shared__ __align__(16) i32 stage_i[];  // mixed: cap * DP_LS ints; FP64: cap * DP_D doubles + cap types
  const int tid = threadIdx.x;
#if DP_NTYPES > 1
  __shared__ DpPC spc[DP_NTYPES * DP_NTYPES];  // filled per tile (the tile's potential)
#else
  DpPC* spc = nullptr;
#endif
  bool trip = false;  // some particle of this lane moved beyond half the skin
  const i64 chunk = (a.ntiles + 7) / 8;
  const i64 xcd = blockIdx.x & 7;
  const i64 cend = (xcd + 1) * chunk < a.ntiles ? (xcd + 1) * chunk : a.ntiles;
  for (i64 j = 0;; ++j) {
    const i64 t = dp_tile_of(j, a.ntiles);
    if (t >= cend) break;
    // (mixed mode: the LDS path reads records by byte-offset entries; with a
    // sentinel, one record of the area is the sentinel's)
    dp_tile_load<DP_D>(S, a.tiletab + t * (i64)a.tstride, a.cap - DP_SENT, DP_ESH != 0);
    const DpRep& P = a.reps[__builtin_amdgcn_readfirstlane(S.rep)];  // uniform: scalar loads
    const int nh = P.nhalo;
    const bool fits = S.fits != 0;
    // rows of the halo: its cells along the fastest dimension (consecutive
    // halo indices, so consecutive staged records and, within each cell,
    // consecutive slots)
    const int rl = P.Th[P.order == 0 ? 0 : DP_D - 1];
    const int nrow = nh / rl;
#if DP_MODE == 1 && (DP_D == 2 || DP_D == 3)
    if (fits) {
      // row by row: wavefront w takes rows w, w + NW, ..; lane l copies
      // records l, l + 64, .. of each (its halo cell from the row's few cell
      // boundaries, no search over the halo), two loads issued before the
      // first one is waited for
      constexpr int NW = DP_TILE_THREADS / 64;
      const int lane = tid & 63;
      const int wave = tid >> 6;
      for (int row = wave; row < nrow; row += NW) {
        const int u0 = row * rl;
        const int b0 = S.hbase[u0], b1 = S.hbase[u0 + rl];
        for (int i0 = b0 + lane; i0 < b1; i0 += 128) {
          DpRec r[2];
          int cu[2];  // halo cell of the record (-1: none)
#pragma unroll
          for (int q = 0; q < 2; ++q) {
            const int i = i0 + 64 * q;
            cu[q] = -1;
            if (i < b1) {
              int u = u0;
              for (int k = 1; k < rl; ++k) u += (i >= S.hbase[u0 + k]) ? 1 : 0;
              r[q] = dp_ld_lrec(a.xq + ((i64)S.hstart[u] + (i - S.hbase[u])) * DP_RS);
              cu[q] = u;
            }
          }
#pragma unroll
          for (int q = 0; q < 2; ++q) {
            if (cu[q] >= 0) {
#pragma unroll
              for (int d = 0; d < DP_D; ++d)
                r[q].v[d] = (i32)((u32)r[q].v[d] + dp_halo_shift<DP_D>(S, P, cu[q], d));
              dp_st_lrec(stage_i + (i0 + 64 * q) * DP_LS, r[q]);
            }
          }
        }
      }
    }
#elif DP_MODE == 1
    if (fits) {
      // 1-D (one long row) and 4-D (rows of a few records): DP_SB records per
      // thread at a time, their loads all issued before the first one is
      // waited for
      for (int b0 = tid; b0 < S.nrec; b0 += DP_SB * DP_TILE_THREADS) {
        DpRec r[DP_SB];
        u32 sh[DP_SB][DP_D];
#pragma unroll
        for (int q = 0; q < DP_SB; ++q) {
          const int idx = b0 + q * DP_TILE_THREADS;
          if (idx < S.nrec) {
            const int u = dp_halo_of<DP_D>(S, nh, idx);
            r[q] = dp_ld_lrec(a.xq + ((i64)S.hstart[u] + (idx - S.hbase[u])) * DP_RS);
#pragma unroll
            for (int d = 0; d < DP_D; ++d) sh[q][d] = dp_halo_shift<DP_D>(S, P, u, d);
          }
        }
#pragma unroll
        for (int q = 0; q < DP_SB; ++q) {
          const int idx = b0 + q * DP_TILE_THREADS;
          if (idx < S.nrec) {
#pragma unroll
            for (int d = 0; d < DP_D; ++d) r[q].v[d] = (i32)((u32)r[q].v[d] + sh[q][d]);
            dp_st_lrec(stage_i + idx * DP_LS, r[q]);
          }
        }
      }
    }
#else
    if (fits) {
      double* sxd = (double*)stage_i;
      i32* stype = (i32*)(sxd + (size_t)a.cap * DP_D);
      for (int idx = tid; idx < S.nrec; idx += blockDim.x) {
        const i64 slot = dp_slot_of(S, nh, idx);
        double t[DP_D];
        dp_ld_vec(a.x + slot * DP_D, t);
#pragma unroll
        for (int d = 0; d < DP_D; ++d) sxd[idx * DP_D + d] = t[d];
        if (DP_NTYPES > 1) stype[idx] = a.type[slot];
      }
    }
#endif
#if DP_MODE == 1 && DP_SENT
    if (fits && tid == 0) {
      // the sentinel record (staged index nrec, the rows' padding entry):
      // 2^31 fixed-point units from the home box's centre along every
      // dimension, so that its exact integer displacement from any home
      // particle lies beyond the cut (the host checks this bound per plan)
      DpRec r;
#pragma unroll
      for (int q = 0; q < DP_RS; ++q) r.v[q] = 0;
#pragma unroll
      for (int d = 0; d < DP_D; ++d) {
        const u32 c = (P.nc[d] == 1) ? 0u : (u32)(((i64)(P.T[d] - 1) * (i64)P.Wq[d]) / 2);
        r.v[d] = (i32)(c + 0x80000000u);
      }
      dp_st_lrec(stage_i + S.nrec * DP_LS, r);
    }
#endif
#if DP_NTYPES > 1
    DP_POT_CALL(P.pot, dp_fill_spc(spc, a.tab[P.pot]));
#endif
    __syncthreads();
    bool trip_t = false;  // a particle of this tile (replica S.rep) moved beyond half the skin
    // replicas with a single cell along some dimension take the copy with the
    // minimum image in the pair loop; the others (usually all) skip it
    if (DP_CMASK != 0 && P.collapsed_mask != 0) {
      if (fits) {
        DP_POT_CALL(P.pot, dp_tile_particles<KIND, true, DP_CMASK>(a, S, P, stage_i, spc, trip_t));
      } else {
        DP_POT_CALL(P.pot, dp_tile_particles<KIND, false, DP_CMASK>(a, S, P, stage_i, spc, trip_t));
      }
    } else if (fits) {
      DP_POT_CALL(P.pot, dp_tile_particles<KIND, true, 0>(a, S, P, stage_i, spc, trip_t));
    } else {
      DP_POT_CALL(P.pot, dp_tile_particles<KIND, false, 0>(a, S, P, stage_i, spc, trip_t));
    }
    if (KIND == 2 || KIND == 3) {
      dp_raise_replica(trip_t, a.rflag, S.rep);
      trip |= trip_t;
    }
    __syncthreads();
  }
  if (KIND == 2 || KIND == 3) dp_raise_rebuild(trip, dp_need_flag(a.ctl, a.blk_pos));
}

// four workgroups (16 wavefronts) per compute unit: at most 128 VGPRs
extern "C" __global__ void __launch_bounds__(DP_TILE_THREADS, 4) dp_tile_force(DpPairArgs a) { dp_tile_run<0>(a); }
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
// The wavefront's slice [w0 * D, (w0 + nv) * D) of an [N][D] array from each
// lane's val[0..D-1], written through LDS so that every store instruction
// covers consecutive addresses (lane e of store k writes element e + 64 k):
// whole cache lines instead of a third of each line per instruction.  Called
// by every lane of the wavefront.
__device__ __forceinline__ void dp_wave_store(double* dst, i64 w0, int nv, const double (&val)[DP_D], double* buf) {
  const int lane = (int)(threadIdx.x & 63);
#pragma unroll
  for (int d = 0; d < DP_D; ++d) buf[lane * DP_D + d] = val[d];
  __builtin_amdgcn_wave_barrier();
#pragma unroll
  for (int k = 0; k < DP_D; ++k) {
    const int e = lane + 64 * k;
    const double v = buf[e];
    if (e < nv * DP_D) dst[w0 * DP_D + e] = v;
  }
  __builtin_amdgcn_wave_barrier();
}

template <int INTEG>
__device__ __forceinline__ void dp_integrate_body(const DpPairArgs& a) {
  __shared__ double wbuf[4][64 * DP_D];  // one slice per wavefront (256 threads)
  // a skipped step of a step block (uniform) returns before loading anything
  if (!dp_blk_go(a.ctl, a.blk_pos)) return;
  const i64 i = (i64)blockIdx.x * blockDim.x + threadIdx.x;
  const bool valid = i < a.N;
  const i64 ii = valid ? i : 0;
  // the particle's state is loaded first, so that the loads are in flight
  // while the replica parameters and the noise are computed
  double x0[DP_D], v0[DP_D], f0[DP_D], b0[DP_D], xr0[DP_D];
  dp_ld_vec(a.x + ii * DP_D, x0);
  dp_ld_vec(a.v + ii * DP_D, v0);
  dp_ld_vec(a.F + ii * DP_D, f0);
#pragma unroll
  for (int d = 0; d < DP_D; ++d) b0[d] = xr0[d] = 0.0;
  if (INTEG == 1) dp_ld_vec(a.beta + ii * DP_D, b0);
  if (a.wrap) dp_ld_vec(a.xref + ii * DP_D, xr0);
  const double m = a.unit_mass ? 1.0 : a.mass[ii];
  const u32 pid = (INTEG == 1) ? (u32)a.orig[ii] : 0u;
#if DP_MODE == 1
  const u32 pack = a.wrap ? (u32)a.cellpack[ii] : 0u;
  DpRec rec;  // the record; its slots after the coordinates (type; 0 with one type) carry over
#pragma unroll
  for (int q = 0; q < DP_RS; ++q) rec.v[q] = 0;
  if (a.wrap && DP_NTYPES > 1) rec = dp_ld_rec(a.xq + ii * DP_RS);
#endif
  const int r = (a.R == 1) ? 0 : a.rep[ii];
  const DpRep& P = a.reps[r];
  double sdisp = 0.0;
  double xn[DP_D], v1[DP_D], nb[DP_D];
#pragma unroll
  for (int d = 0; d < DP_D; ++d) xn[d] = v1[d] = nb[d] = 0.0;
  if (valid) {
#pragma clang fp contract(off)
    const double dt = P.dt;
    if (INTEG == 1) {
      double sq, ca, cb;
      dp_gjf_coeffs(P, m, sq, ca, cb);
      // normal numbers in pairs: one Philox block and Box-Muller per two dimensions
      double z[DP_D];
      if (a.xi) {
#pragma unroll
        for (int d = 0; d < DP_D; ++d) z[d] = a.xi[(i64)pid * DP_D + d];
      } else {
        const u64 step = dp_step_index(a.ctl, a.blk_pos);
#pragma unroll
        for (int k = 0; 2 * k < DP_D; ++k) {
          double z0, z1;
          dp_normal2(P.seed, step, (u32)(pid - (u32)P.p0), k, z0, z1);  // replica seed, id in replica
          z[2 * k] = z0;
          if (2 * k + 1 < DP_D) z[2 * k + 1] = z1;
        }
      }
#pragma unroll
      for (int d = 0; d < DP_D; ++d) {
        nb[d] = sq * z[d];
        v1[d] = dp_gjf_v(ca, cb, dt, m, v0[d], f0[d], b0[d], nb[d]);
        xn[d] = x0[d] + v1[d] * dt;
      }
    } else {
#pragma unroll
      for (int d = 0; d < DP_D; ++d) {
        v1[d] = dp_lf_v(dt, m, v0[d], f0[d]);
        xn[d] = x0[d] + dt * v1[d];
      }
    }
#pragma unroll
    for (int d = 0; d < DP_D; ++d) {
      if (a.wrap) {
        const double sh = dp_wrap_shift(xn[d], P.L[d]);
        if (!(xn[d] > 0.0 && xn[d] < P.L[d])) {
          a.img[ii * DP_D + d] += (i32)sh;
          xn[d] = xn[d] - sh * P.L[d];
        }
        double tq = dp_sub(xn[d], xr0[d]);
        tq = dp_minimage_gt(tq, P.L[d], P.halfL[d]);
        sdisp = (d == 0) ? dp_mul(tq, tq) : dp_add(sdisp, dp_mul(tq, tq));
      }
    }
#if DP_MODE == 1
    if (a.wrap) {
#pragma unroll
      for (int d = 0; d < DP_D; ++d) {
        const u32 c = (pack >> (DP_PACK_BITS * d)) & ((1u << DP_PACK_BITS) - 1u);
        const double cen = ((double)c + 0.5) * P.w[d];
        rec.v[d] = (i32)rint(dp_minimage_gt(xn[d] - cen, P.L[d], P.halfL[d]) * a.tab->qinv);
      }
      dp_st_rec(a.xq_out + ii * DP_RS, rec);
    }
#endif
  }
  // the wavefront's velocities, betas and positions, whole lines at a time
  const i64 w0 = i - (threadIdx.x & 63);
  const int nv = (int)((a.N - w0) < 64 ? (a.N - w0) : 64);
  double* buf = wbuf[threadIdx.x >> 6];
  dp_wave_store(a.v, w0, nv, v1, buf);
  if (INTEG == 1) dp_wave_store(a.beta, w0, nv, nb, buf);
  dp_wave_store(a.x_out, w0, nv, xn, buf);
  if (a.wrap) {
    const bool trip = valid && dp_skin_trips(sdisp, P.skin_half_sq);
    dp_raise_replica(trip, a.rflag, r);
    dp_raise_rebuild(trip, dp_need_flag(a.ctl, a.blk_pos));
  }
}

extern "C" __global__ void __launch_bounds__(256, 5) dp_integrate_gjf(DpPairArgs a) { dp_integrate_body<1>(a); }
extern "C" __global__ void __launch_bounds__(256, 5) dp_integrate_lf(DpPairArgs a) { dp_integrate_body<2>(a); }
