// dp_lib.hip -- host side of libdompap_hip.so: context and device memory,
// hiprtc compilation of the pair kernels, the per-step launch sequence (with
// hipGraph capture of step blocks), and the C-ABI declared in
// include/dompap_hip.h.
#include "dp_kernels.h"

#include <hip/hiprtc.h>

#include <algorithm>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <map>
#include <mutex>
#include <string>
#include <vector>

#include "dp_embedded.inc"  // dp_src_device_h, dp_src_pair (generated at build time)
#include "../../include/dompap_hip.h"

#define DP_API extern "C" __attribute__((visibility("default")))

// ---------------------------------------------------------------------------
// Device scratch of the allocation-free entry point (dp_rdf_counts_noalloc):
// the bin thresholds, the histogram and the per-workgroup displacement
// maxima of one call, at most DP_RDF_MAXBINS bins.  Each device holds its own
// copy until the process ends; calls on one device share it (overlapping
// calls on one device race).  Host code never passes the symbol itself to a
// kernel: it uses the device address that hipGetSymbolAddress returns, kept
// per device.
// ---------------------------------------------------------------------------
#define DP_NA_MAX_EDGES (DP_RDF_MAXBINS + 1)
#define DP_NA_MAX_BLOCKS 512  // grid_for() caps the displacement grid here
struct DpNoAllocLayout {
  double thr[DP_NA_MAX_EDGES];
  unsigned long long hist[DP_RDF_MAXBINS];
  double block_max[DP_NA_MAX_BLOCKS];
};
static __device__ __align__(256) unsigned char scratch[sizeof(DpNoAllocLayout)];

#define DP_MAX_DEVICES 64
namespace {
std::mutex g_scratch_mu;
void* g_scratch_ptr[DP_MAX_DEVICES] = {};

// Device address of `scratch` on the current device `dev` (looked up once per
// device, with that device current).
int scratch_address(int dev, void** out, std::string& err) {
  if (dev < 0 || dev >= DP_MAX_DEVICES) {
    err = "device index beyond the scratch table";
    return DP_ERR_ARG;
  }
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  if (!g_scratch_ptr[dev]) {
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p) {
      err = std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e);
      return DP_ERR_HIP;
    }
    g_scratch_ptr[dev] = p;
  }
  *out = g_scratch_ptr[dev];
  return DP_OK;
}

__global__ void k_fill_u64(unsigned long long* p, int n) {
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) p[i] = 0ull;
}
}  // namespace

namespace {

std::string g_create_error;

int rs_for_dim(int D) {
  int rs = 2;
  while (rs < D + 1) rs *= 2;
  return rs;
}

// Entry shift of the pair kernels' staged records: mixed precision, log2 of
// the record's bytes (the coordinates, + the type with several types, in a
// power of two of ints: DP_LS of the pair source), so that an entry is a
// byte offset; FP64 mode, 0.
int staged_esh(int D, int mode, int ntypes) {
  if (mode != DP_MIXED) return 0;
  int ls = 1;
  while (ls < D + (ntypes > 1 ? 1 : 0)) ls *= 2;
  int esh = 2;
  while ((1 << esh) < 4 * ls) ++esh;
  return esh;
}

template <typename T>
void dfree(T*& p) {
  if (p) (void)hipFree((void*)p);
  p = nullptr;
}

// Exact thresholds (monotone predicates over non-negative doubles).
double bits2d(uint64_t b) { double d; std::memcpy(&d, &b, 8); return d; }
uint64_t d2bits(double d) { uint64_t b; std::memcpy(&b, &d, 8); return b; }

// Smallest s with !(sqrt(s) < cutoff): then  s < t  <=>  sqrt(s) < cutoff.
double nl_threshold(double cutoff) {
  auto inside = [&](double s) { return std::sqrt(s) < cutoff; };
  if (!inside(0.0)) return 0.0;
  const double inf = INFINITY;
  if (inside(inf)) return inf;
  uint64_t lo = 0, hi = d2bits(inf);
  while (hi - lo > 1) {
    const uint64_t mid = lo + (hi - lo) / 2;
    if (inside(bits2d(mid))) lo = mid; else hi = mid;
  }
  return bits2d(hi);
}

// Largest s with !(sqrt(s)/sigma > r_cut): then  s > t  <=>  sqrt(s)/sigma > r_cut.
double cut_threshold(double sigma, double r_cut) {
  auto outside = [&](double s) { return std::sqrt(s) / sigma > r_cut; };
  if (outside(0.0)) return -1.0;
  const double inf = INFINITY;
  if (!outside(inf)) return inf;
  uint64_t lo = 0, hi = d2bits(inf);
  while (hi - lo > 1) {
    const uint64_t mid = lo + (hi - lo) / 2;
    if (outside(bits2d(mid))) hi = mid; else lo = mid;
  }
  return bits2d(lo);
}

struct CodeCache {
  std::mutex mu;
  std::map<std::string, std::vector<char>> code;
};
CodeCache& code_cache() {
  static CodeCache c;
  return c;
}


struct PairConfig {
  int D, mode, ntypes, even, rs, tpp, nlw, ob, pairprec;
  int cmask;  // dimensions with a single cell in some replica
  int esh = 0, sent = 0;  // entry = staged index << esh; rows padded with the sentinel's entry
  int hmax = 0, omax = 0;  // LDS tile-table capacity: largest halo / home cell count of the plan (0: the limits)
};

// Stencil-index bits of a neighbour entry (3^D stencil cells).
int stencil_bits(int D) {
  int n3 = 1;
  for (int d = 0; d < D; ++d) n3 *= 3;
  int b = 0;
  while ((1 << b) < n3) ++b;
  return b;
}

// Potential-dependent section of the pair-kernel source (between the
// markers), compiled once per potential inside namespace dp_p<k>.
void split_pair_source(std::string& pre, std::string& pot, std::string& post) {
  const std::string all = dp_src_pair;
  const std::string b = "//@@POTENTIAL_BEGIN", e = "//@@POTENTIAL_END";
  const size_t i = all.find(b), j = all.find(e);
  pre = all.substr(0, i);
  pot = all.substr(i, j - i);
  post = all.substr(j);
}

// The pair kernels' source for potentials srcs[k] (generated device
// functions; evens[k] = 1 when they take s = x^2), replica r using potential
// DpRep::pot.
std::string pair_source(const PairConfig& c, const std::vector<std::string>& srcs, const std::vector<int>& evens,
                        const std::vector<int>& folds) {
  std::string src;
  char buf[512];
  snprintf(buf, sizeof(buf),
           "#define DP_D %d\n#define DP_MODE %d\n#define DP_NTYPES %d\n#define DP_NPOT %d\n#define DP_RS %d\n"
           "#define DP_TPP %d\n#define DP_NLW %d\n#define DP_OB %d\n#define DP_PAIRPREC %d\n#define DP_CMASK %d\n"
           "#define DP_ESH %d\n#define DP_SENT %d\n",
           c.D, c.mode, c.ntypes, (int)srcs.size(), c.rs, c.tpp, c.nlw, c.ob, c.pairprec, c.cmask, c.esh, c.sent);
  src += buf;
  if (c.hmax > 0 && c.omax > 0) {
    snprintf(buf, sizeof(buf), "#define DP_TS_HALO %d\n#define DP_TS_HOME %d\n", c.hmax, c.omax);
    src += buf;
  }
  src += "#include \"dp_device.h\"\n";
  std::string pre, pot, post;
  split_pair_source(pre, pot, post);
  src += pre;
  std::string call = "#define DP_POT_CALL(pot, ...) switch (pot) {";
  for (size_t k = 0; k < srcs.size(); ++k) {
    const int fold = (k < folds.size()) ? folds[k] : 0;
    snprintf(buf, sizeof(buf), "namespace dp_p%d {\n#define DP_EVEN %d\n#define DP_FOLD %d\n", (int)k,
             evens[k] ? 1 : 0, fold);
    src += buf;
    src += srcs[k];
    src += "\n";
    src += pot;
    src += "#undef DP_EVEN\n#undef DP_FOLD\n}\n";
    snprintf(buf, sizeof(buf), " case %d: { using namespace dp_p%d; __VA_ARGS__; } break;", (int)k, (int)k);
    call += buf;
  }
  call += " }\n";
  src += call;
  src += post;
  return src;
}

std::string pair_source(const PairConfig& c, const char* device_src, int fold) {
  return pair_source(c, std::vector<std::string>{device_src}, std::vector<int>{c.even}, std::vector<int>{fold});
}

// The "// dp-gsk: n e_0 c_0 .." line of a generated potential source: f(x)/x
// = sum_k c_k s^-e_k with s = x^2 (at most four terms); false if absent.
bool parse_gsk(const std::string& src, int& n, int* e, double* cf) {
  const size_t at = src.find("// dp-gsk: ");
  if (at == std::string::npos) return false;
  const char* p = src.c_str() + at + 11;
  char* end = nullptr;
  const long nn = std::strtol(p, &end, 10);
  if (end == p || nn < 1 || nn > 4) return false;
  p = end;
  for (int k = 0; k < (int)nn; ++k) {
    const long ek = std::strtol(p, &end, 10);
    if (end == p || ek < 0 || ek > 64) return false;
    p = end;
    const double ck = std::strtod(p, &end);
    if (end == p || !std::isfinite(ck)) return false;
    p = end;
    e[k] = (int)ek;
    cf[k] = ck;
  }
  n = (int)nn;
  return true;
}

// hiprtc compile for gfx950 (no device needed), memoised per process.
int compile_cached(const std::string& src, std::vector<char>& code, std::string& log) {
  {
    std::lock_guard<std::mutex> lk(code_cache().mu);
    auto it = code_cache().code.find(src);
    if (it != code_cache().code.end()) {
      code = it->second;
      return DP_OK;
    }
  }
  hiprtcProgram prog;
  const char* hdrs[] = {dp_src_device_h};
  const char* names[] = {"dp_device.h"};
  if (hiprtcCreateProgram(&prog, src.c_str(), "dp_pair.hip", 1, hdrs, names) != HIPRTC_SUCCESS) {
    log = "hiprtcCreateProgram failed";
    return DP_ERR_RTC;
  }
  // packed-FP32 instructions are disabled for all HIP code of this project
  const char* opts[] = {"--offload-arch=gfx950", "-O3", "-std=c++17", "-ffp-contract=fast-honor-pragmas",
                        "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops"};
  hiprtcResult rr = hiprtcCompileProgram(prog, (int)(sizeof(opts) / sizeof(opts[0])), opts);
  size_t ls = 0;
  hiprtcGetProgramLogSize(prog, &ls);
  log.assign(ls, '\0');
  if (ls) hiprtcGetProgramLog(prog, &log[0]);
  if (rr != HIPRTC_SUCCESS) {
    hiprtcDestroyProgram(&prog);
    return DP_ERR_RTC;
  }
  size_t cs = 0;
  hiprtcGetCodeSize(prog, &cs);
  code.resize(cs);
  hiprtcGetCode(prog, code.data());
  hiprtcDestroyProgram(&prog);
  std::lock_guard<std::mutex> lk(code_cache().mu);
  code_cache().code[src] = code;
  return DP_OK;
}

}  // namespace

// Timing events of one step: before the force kernel, between the force and
// integrate kernels, after the integrate kernel.
struct DpEv3 {
  hipEvent_t e[3];
};

struct dp_ctx {
  int dev = 0;
  int ncu = 256;  // compute units of the device
  int lds_cu = 160 * 1024;  // LDS bytes per compute unit
  int integ_static_lds = 0;  // static LDS of the integrate kernels
  hipStream_t st = nullptr;
  int D = 3, mode = DP_FP64, R = 1, rs = 4;
  i64 N = 0;
  int K = 0, order = DP_ORDER_X_FASTEST;
  i64 Ctot = 0;
  std::string err;

  // host mirrors of parameters
  std::vector<i64> offsets;
  std::vector<double> box, cutoff, skin, dt, T, tau;
  bool have_box = false, have_nl_params = false, have_integ = false;
  std::vector<DpRep> reps_h;
  std::vector<DpPairTab> tabs_h;  // one per potential
  std::vector<i32> rep_pot;       // potential of each replica (empty: all use potential 0)
  int ntypes = 1;
  bool have_pot = false;
  bool reps_dirty = true, tab_dirty = true;

  // device memory
  DpRep* reps_d = nullptr;
  DpPairTab* tab_d = nullptr;
  DpCtl* ctl_d = nullptr;
  DpCtl* ctl_h = nullptr;  // pinned
  i64* nlupd_d = nullptr;
  i32* rflag = nullptr;  // [2R]: need[r] (skin trigger of replica r), rebuild[r] (current step)
  double* x[2] = {nullptr, nullptr};
  i32* xq[2] = {nullptr, nullptr};
  double *v = nullptr, *beta = nullptr, *mass = nullptr, *xref = nullptr, *F = nullptr;
  double *E = nullptr, *W = nullptr, *LAP = nullptr, *xi = nullptr, *red = nullptr;
  i32 *img = nullptr, *type = nullptr, *orig = nullptr, *rep = nullptr, *cellpack = nullptr;
  i32 *cell_of = nullptr, *rank = nullptr, *tmp = nullptr, *perm = nullptr, *nlcnt = nullptr;
  i32 *ccount = nullptr, *cstart = nullptr;
  void* nl = nullptr;
  size_t nl_bytes = 0;
  // tile / entry-format plan (refresh_reps)
  int nlw = 16, ob = 11, tpp = 8, qbits = 24, cap = 1024, cap_emit = 1024;
  int efmt = 0;  // neighbour-entry format: esh | sent << 8 (dp_efmt_hlim)
  int emit_lanes = 0;  // 1: the emit takes one lane per home particle (cells of a few particles)
  int nlw_min = 16;    // 32 after a tile's halo outgrew 16-bit entries (kept for later builds)
  bool unit_mass = true;  // every mass is 1.0 (as at creation)
  bool last_cell_overflow = false;  // the last check_ctl failed on a halo beyond the entry width
  // tuning (0 = automatic)
  int tune_tpp = 0, tune_tile = 0, tune_fused = -1;
  double cell_scale = 1.0;  // cells at least cell_scale x cutoff wide (tuning; >= 1)
  int fused = 0;  // 1: fused force + integrate kernel; 0: force kernel, then integrate kernel
  int cmask = 0;  // OR of the replicas' collapsed-dimension masks
  i64 ntiles = 0;
  int tile_grid = 8;
  size_t lds_pair = 0, lds_emit = 0;
  // potential (compiled lazily once the plan is known)
  std::vector<std::string> pot_srcs;  // generated device functions, one per potential
  std::vector<int> pot_evens;
  int pair_prec = 64;
  std::string mod_key;
  char* scratch = nullptr;
  size_t scratch_bytes = 0;
  char* stage = nullptr;
  int cur = 0;

  // pair kernels
  hipModule_t mod = nullptr;
  hipFunction_t f_force = nullptr, f_obs = nullptr, f_gjf = nullptr, f_lf = nullptr;
  hipFunction_t f_int_gjf = nullptr, f_int_lf = nullptr;

  // state flags
  bool nl_built = false;
  i32* tiletab = nullptr;  // per-tile tables written at every build
  i32* todo = nullptr;     // [ntiles] the emit's tile list (batches)
  int2* plist = nullptr;   // batches: particle blocks of the rebuilding replicas (k_rebuild_list)
  i32* rlist = nullptr;    //   and the rebuilding replicas
  i64 plist_n = 0;
  i64 todo_n = 0;
  int tstride = 0;
  size_t tiletab_bytes = 0;
  bool nl_valid = true;                 // false when the list buffer was reallocated
  int rdf_path = -1;                    // path of the last dp_rdf_counts: 1 cell tiles, 0 all pairs
  void* na_scratch = nullptr;           // this device's address of `scratch` (set by dp_create)
  std::vector<i64> cells_sig, tiles_sig;  // geometry the list was built for
  bool pos_dirty = true;
  u64 seed = 0x5eed0d0bull;
  std::vector<u64> rep_seeds;  // per replica; empty: every replica uses `seed`

  // graphs
  int graph_block = 16;
  int block_mode = -1;  // 0: rebuild once per block, 1: before every step (gated per replica), -1: 1 iff R > 1
  hipGraphExec_t gexec[2][2] = {{nullptr, nullptr}, {nullptr, nullptr}};  // [timed][starting parity]
  int gexec_integ[2][2] = {{0, 0}, {0, 0}};
  std::vector<DpEv3> blk_ev;            // timed step blocks: events of step p (recorded by the graph)
  bool graph_events_ok = true;          // events recorded inside graphs work on this runtime
  bool capturing = false;               // a step block is being captured (events become graph nodes)
  bool graph_dirty = true;

  // largest staged halo (records) measured by an emit so far: raises the
  // LDS record areas in inhomogeneous systems; the plan's mean halo
  // occupancy and staged record size
  int halo_seen = 0;
  double halo_occ = 0.0;
  size_t rec_bytes = 0;
  // kernel timing
  bool timing = false;
  int timing_every = 1;   // step blocks: events in one block of every timing_every
  int64_t timing_block = 0;  // step blocks launched since timing was enabled
  std::vector<DpEv3> ev_pending;
  std::vector<DpEv3> ev_free;
  double time_ms = 0.0, time_force_ms = 0.0, time_integ_ms = 0.0;
  i64 time_launches = 0;

  int fail(int code, const char* fmt, ...) {
    char buf[2048];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof(buf), fmt, ap);
    va_end(ap);
    err = buf;
    return code;
  }
};

#define HIPCHK(call)                                                                    \
  do {                                                                                  \
    hipError_t e_ = (call);                                                             \
    if (e_ != hipSuccess) return ctx->fail(DP_ERR_HIP, "%s: %s", #call, hipGetErrorString(e_)); \
  } while (0)

#define CHK(expr)               \
  do {                          \
    int rc_ = (expr);           \
    if (rc_ != DP_OK) return rc_; \
  } while (0)

namespace {

// Grid of the grid-stride pipeline kernels: at most 512 workgroups, so that
// a launch skipped by its rebuild gate retires quickly.
int grid_for(i64 n, int cap = 512) {
  i64 b = (n + DP_BLOCK - 1) / DP_BLOCK;
  if (b < 1) b = 1;
  if (b > cap) b = cap;
  return (int)b;
}

template <typename T>
int dalloc(dp_ctx* ctx, T*& p, size_t count) {
  dfree(p);
  if (count == 0) count = 1;
  HIPCHK(hipMalloc((void**)&p, count * sizeof(T)));
  return DP_OK;
}

int release_graph(dp_ctx* ctx) {
  for (auto& row : ctx->gexec)
    for (auto& g : row) {
      if (g) (void)hipGraphExecDestroy(g);
      g = nullptr;
    }
  ctx->graph_dirty = true;
  return DP_OK;
}

// Volume of the unit D-ball.
double ball_volume(int D) {
  const double pi = 3.14159265358979323846;
  switch (D) {
    case 1: return 2.0;
    case 2: return pi;
    case 3: return 4.0 * pi / 3.0;
    case 4: return pi * pi / 2.0;
  }
  return 1.0;
}

// Cells, tiles, fixed-point scale, entry format and LDS capacity from the
// boxes, cutoffs and particle counts.
//  - cells: nc = floor(L / cutoff) per dimension (1 if fewer than 3 fit);
//  - lanes per particle (tpp) from the expected list length;
//  - tiles: enough home cells that two passes of the workgroup's lane groups
//    are about full (one pass for systems too small to give 8 such tiles per
//    compute unit);
//  - fixed-point scale: coordinates relative to a tile's first home cell
//    centre stay below 2^30 units;
//  - entries: index of the neighbour in its tile's staged halo, 16-bit when
//    halos are small enough, else 32-bit.
// A new plan over the same cells re-emits the existing list (from xref)
// instead of rebuilding it, so rebuild times do not depend on the plan.
// LDS record areas of the pair kernels and the emit: the mean halo plus
// three standard deviations of a Poisson count (at most 1.5x the mean) and
// 64, at most 48 KB; the rare fuller tiles take the global-memory paths.
// (4-D tiles: 2,027 records instead of 2,817, which brings the force kernel
// to 4 workgroups per CU.)  A larger halo measured by an emit (inhomogeneous
// systems: the dense tiles) raises both areas as far as four workgroups per
// compute unit allow.  Launch sizes only: the list and the plan stay valid.
void size_lds(dp_ctx* ctx) {
  const double h = ctx->halo_occ;
  const i64 stat = (i64)std::ceil(std::min(1.5 * h, h + 3.0 * std::sqrt(h))) + 64;
  const i64 want = (i64)ctx->halo_seen + 32;
  // pair kernels: 40 KB of LDS per workgroup at four per CU, ~2 KB of it the
  // tile tables
  i64 cap = stat;
  if (want > cap) cap = std::max<i64>(cap, std::min<i64>(want, (i64)((38 * 1024) / ctx->rec_bytes)));
  cap = std::min<i64>(cap, (i64)(48 * 1024 / ctx->rec_bytes));
  cap = std::max<i64>(cap, 64);
  if ((int)cap != ctx->cap) release_graph(ctx);  // captured launches carry the LDS size
  ctx->cap = (int)cap;
  ctx->lds_pair = (size_t)cap * ctx->rec_bytes;
  // emit: 18 bytes per record (an int4 record and a u16 candidate-table
  // entry), ~7.3 KB of tile tables; four emit workgroups fit a compute unit
  // up to 1,848 records.  Where that bound still covers the mean halo plus
  // three standard deviations, the planned area drops its 64-record margin
  // to stay within it (3-D LJ at N = 2^20 on 34 cells per side: 1,897
  // records by the rule, three emit workgroups per compute unit and 730 us
  // per rebuild, against 605 us at 1,848).  Fuller halos keep the rule's
  // area: with the mean near the bound (4-D, replica batches) most tiles
  // would take the global path (256 replicas 2.44e9 -> 1.90e9).
  const i64 ce4 = (i64)((40 * 1024 - 7680) / 18);
  i64 ce = (stat > ce4 && stat - 64 <= ce4) ? ce4 : stat;
  if (want > ce) ce = std::max<i64>(ce, std::min<i64>(want, ce4));
  ce = std::min<i64>(ce, (i64)(48 * 1024 / 16));
  ce = std::max<i64>(ce, 64);
  if ((int)ce != ctx->cap_emit) release_graph(ctx);
  ctx->cap_emit = (int)ce;
  ctx->lds_emit = (size_t)ce * 16 + ((size_t)ce * 2 + 15) / 16 * 16;
}

int launch_records(dp_ctx* ctx);
int launch_reemit(dp_ctx* ctx);
int refresh_reps(dp_ctx* ctx) {
  if (!ctx->reps_dirty) return DP_OK;
  if (!ctx->have_box) return ctx->fail(DP_ERR_STATE, "dp_set_replicas has not been called");
  if (!ctx->have_nl_params) return ctx->fail(DP_ERR_STATE, "dp_set_neighbor_params has not been called");
  const int D = ctx->D;
  const int bits = (D == 1) ? 31 : 32 / D;
  const i64 maxc = (D == 1) ? (i64)0x7fffffff : ((i64)1 << bits) - 1;
  std::vector<DpRep> reps((size_t)ctx->R);
  // expected list length -> lanes per particle
  double kbar_max = 0.0;
  for (int r = 0; r < ctx->R; ++r) {
    double vol = 1.0;
    for (int d = 0; d < D; ++d) vol *= ctx->box[(size_t)r * D + d];
    const double n = (double)(ctx->offsets[r + 1] - ctx->offsets[r]);
    const double cut = std::isfinite(ctx->cutoff[r]) ? ctx->cutoff[r] : 1e300;
    const double kb = std::min(n, n / vol * ball_volume(D) * std::pow(std::max(cut, 0.0), D));
    kbar_max = std::max(kbar_max, kb);
  }
  // measured on MI355X (3D LJ, ~95 entries per row): 2 lanes per particle
  // and the split step are fastest; the fused kernel idles lanes while one
  // lane per dimension integrates.  In mixed mode, one system large enough
  // for two passes of 256 particles per tile and 8 such tiles per compute
  // unit takes one lane per particle (3-D LJ at N = 2^20: force kernel
  // 169.7 -> 164.6 us; 256 replicas of ~5,000 particles -18 % and N = 4,096
  // -20 % with one lane; the FP64 kernel 506 -> 555 us)
  int tpp = 1;
  if (kbar_max >= 160) tpp = 4;
  else if (kbar_max >= 24) tpp = 2;
  if (tpp == 2 && ctx->mode == DP_MIXED && ctx->R == 1 &&
      (double)ctx->N / (2.0 * DP_TILE_THREADS) >= 8.0 * ctx->ncu)
    tpp = 1;
  // small systems: more lanes per particle (up to 16) while one pass of lane
  // groups per tile would leave fewer tiles than compute units, so that a
  // step's work spreads over the device (measured on MI355X: 3-D LJ at
  // N = 4,096, 2 -> 16 lanes per particle: mixed 1.32e8 -> 1.63e8, FP64
  // 9.3e7 -> 1.47e8 particle-steps/s)
  while (tpp < 16 && (double)ctx->N * tpp < (double)DP_TILE_THREADS * ctx->ncu) tpp *= 2;
  if (ctx->tune_tpp > 0) tpp = ctx->tune_tpp;
  ctx->fused = ctx->tune_fused >= 0 ? ctx->tune_fused : 0;
  // home particles per tile: two passes of the workgroup's lane groups
  // (fewer tiles stage fewer halo records per particle; the last pass's
  // leftover particles take the idle wavefronts with more lanes each), or
  // one pass where two would leave fewer than 8 tiles per compute unit
  // (measured on MI355X: 3-D LJ at N = 2^20 +6.6 %, 2-D r^-12 at N = 2^22
  // +8 %; 3-D LJ at N = 4,096 -13 % with two passes)
  const int pass = DP_TILE_THREADS / tpp;
  const bool two = (double)ctx->N / (2.0 * pass) >= 8.0 * ctx->ncu;
  const int target = ctx->tune_tile > 0 ? ctx->tune_tile : (two ? 2 * pass : pass);
  // LDS record capacity of the pair kernels (48 KB of records)
  // staged record: mixed, the coordinates (+ the type with several types) in
  // a power of two of ints (DP_LS of the pair kernels); FP64, D doubles (+ the
  // type with several types)
  int ls = 1;
  while (ls < D + (ctx->ntypes > 1 ? 1 : 0)) ls *= 2;
  const size_t rec_bytes = (ctx->mode == DP_MIXED) ? (size_t)ls * 4
                                                   : (size_t)D * 8 + (ctx->ntypes > 1 ? 4 : 0);
  const i64 lds_rec_max = (i64)(48 * 1024 / rec_bytes);
  i64 c0 = 0, t0 = 0;
  double rmax = 1.0;   // largest |coordinate offset| or displacement, for the fixed-point scale
  double occ_max = 0.0, halo_occ_max = 0.0;
  for (int r = 0; r < ctx->R; ++r) {
    DpRep& P = reps[r];
    std::memset(&P, 0, sizeof(P));
    P.p0 = ctx->offsets[r];
    P.n = ctx->offsets[r + 1] - ctx->offsets[r];
    P.order = ctx->order;
    P.pot = ((int)ctx->rep_pot.size() == ctx->R) ? ctx->rep_pot[r] : 0;
    const double cut = ctx->cutoff[r];
    i64 nc[DP_DMAX];
    i64 ncell = 1;
    for (int d = 0; d < D; ++d) {
      const double L = ctx->box[(size_t)r * D + d];
      double q = (cut > 0.0 && std::isfinite(cut)) ? std::floor(L / (cut * ctx->cell_scale)) : 1.0;
      if (!(q >= 3.0)) q = 1.0;
      if (q > (double)maxc) q = (double)maxc;
      nc[d] = (i64)q;
      ncell *= nc[d];
    }
    // sparse systems: at most ~8 cells per particle
    const i64 cap_cells = std::max<i64>(64, 8 * std::max<i64>(P.n, 1));
    while (ncell > cap_cells) {
      int dm = 0;
      for (int d = 1; d < D; ++d)
        if (nc[d] > nc[dm]) dm = d;
      if (nc[dm] <= 3) break;
      nc[dm] = std::max<i64>(3, nc[dm] / 2);
      ncell = 1;
      for (int d = 0; d < D; ++d) ncell *= nc[d];
    }
    double occ = (double)P.n / (double)ncell;
    // tile shape
    i64 T[DP_DMAX];
    for (int d = 0; d < D; ++d) T[d] = 1;
    for (;;) {
      i64 prodT = 1, prodH = 1;
      for (int d = 0; d < D; ++d) {
        prodT *= T[d];
        prodH *= (nc[d] == 1) ? 1 : T[d] + 2;
      }
      int best = -1;
      for (int d = 0; d < D; ++d) {
        if (nc[d] == 1 || T[d] >= nc[d]) continue;
        if (best < 0 || T[d] < T[best]) best = d;
      }
      if (best < 0) break;
      const i64 nT = prodT / T[best] * (T[best] + 1);
      const i64 nH = prodH / (T[best] + 2) * (T[best] + 3);
      // stop at a full pass of lane groups (mean occupancy; measured on 2-D
      // r^-12 at N = 2^22, 9 x 8 home cells of 233 particles mean beat both
      // 8 x 8 and 9 x 9), the table limits, or a mean halo that would not fit
      // the LDS record area (tiles above the mean by more than the area's
      // slack take the global-memory paths)
      if (occ * (double)nT > (double)target || nT > DP_HOME_MAX || nH > DP_HALO_MAX ||
          occ * (double)nH + 32.0 > (double)lds_rec_max)
        break;
      // beyond one pass only while the halo stays well inside the LDS record
      // area (dense replicas of a batch would otherwise size the area for
      // every tile: 256 mixed replicas -3.5 % with their 8-cell tiles)
      if (occ * (double)nT > (double)pass && occ * (double)nH + 32.0 > 0.6 * (double)lds_rec_max) break;
      T[best] += 1;
    }
    // whole tiles: along a dimension whose cell count is not a multiple of
    // the tile width, the cells are widened (by at most 6 %) to the nearest
    // multiple below, while the tile's mean occupancy stays within its pass
    // target and its halo within the LDS record area.  The last tile of each
    // row was partial (3-D LJ at N = 2^20: 35 cells per side and 18^3 tiles,
    // 16 % of them partial; 34 and 17^3 whole tiles: force kernel 154 ->
    // 141 us, FP64 mode +6.6 %)
    if (ctx->cell_scale == 1.0) {
      i64 nc2[DP_DMAX], ncell2 = 1, prodT = 1, prodH = 1;
      for (int d = 0; d < D; ++d) {
        nc2[d] = nc[d];
        if (nc[d] > 1 && T[d] >= 2 && nc[d] % T[d] != 0) {
          const i64 m = nc[d] - nc[d] % T[d];
          if (m >= 3 && (double)nc[d] <= 1.06 * (double)m) nc2[d] = m;
        }
        ncell2 *= nc2[d];
        prodT *= T[d];
        prodH *= (nc2[d] == 1) ? 1 : T[d] + 2;
      }
      const double occ2 = (double)P.n / (double)ncell2;
      if (ncell2 != ncell && occ2 * (double)prodT <= (double)target &&
          occ2 * (double)prodH + 32.0 <= (double)lds_rec_max) {
        for (int d = 0; d < D; ++d) nc[d] = nc2[d];
        ncell = ncell2;
        occ = occ2;
      }
    }
    i64 ntiles = 1, nhome = 1, nhalo = 1;
    for (int d = 0; d < D; ++d) {
      const double L = ctx->box[(size_t)r * D + d];
      P.nc[d] = (i32)nc[d];
      P.T[d] = (i32)T[d];
      P.nt[d] = (i32)((nc[d] + T[d] - 1) / T[d]);
      P.Th[d] = (nc[d] == 1) ? 1 : (i32)T[d] + 2;
      ntiles *= P.nt[d];
      nhome *= P.T[d];
      nhalo *= P.Th[d];
      P.L[d] = L;
      P.halfL[d] = L / 2;
      P.w[d] = L / (double)nc[d];
      P.inv_w[d] = (double)nc[d] / L;
      // fixed-point range: records hold offsets from their cell centre
      // (|q| <= w/2 + skin between builds); halo-box coordinates and pair
      // displacements are formed in wrapping 32-bit arithmetic, exact while
      // the displacement itself fits (listed pairs: |dq| <= cutoff + 2 skin);
      // a collapsed dimension uses offsets from the box centre
      const double sk = std::max(ctx->skin[r], 0.0);
      if (nc[d] == 1) {
        P.collapsed_mask |= (1 << d);
        rmax = std::max(rmax, L);
      } else {
        rmax = std::max(rmax, P.w[d] + sk);  // also keeps Wq within 2^30
        rmax = std::max(rmax, (std::isfinite(cut) ? cut : 0.0) + 2.0 * sk);
      }
    }
    // a box of home cells holding many more particles than one pass of the
    // workgroup's lane groups (a single-cell replica of hundreds of
    // particles) is split into slices of its home particles, one workgroup
    // each, so that it does not hold up a launch on one compute unit
    {
      const double exp_home = occ * (double)nhome;
      P.nslice = 1;
      if (exp_home > 2.0 * (double)pass)
        P.nslice = (i32)std::min<double>(std::ceil(exp_home / (double)pass), 1024.0);
      ntiles *= P.nslice;
    }
    P.c0 = c0;
    P.ncell = ncell;
    P.t0 = t0;
    P.ntiles = ntiles;
    P.nhome = (i32)nhome;
    P.nhalo = (i32)nhalo;
    c0 += ncell;
    t0 += ntiles;
    occ_max = std::max(occ_max, occ);
    halo_occ_max = std::max(halo_occ_max, occ * (double)nhalo);
    P.nl_thresh = nl_threshold(cut);
    P.skin_half_sq = (0.5 * ctx->skin[r]) * (0.5 * ctx->skin[r]);
    P.dt = ctx->have_integ ? ctx->dt[r] : 0.01;
    P.T_ = ctx->have_integ ? ctx->T[r] : 1.0;
    P.tau = ctx->have_integ ? ctx->tau[r] : 0.1;
    P.seed = ((int)ctx->rep_seeds.size() == ctx->R) ? ctx->rep_seeds[r] : ctx->seed;
    {
      // G-JF coefficients of a unit mass, in integrator.py:17-40's order
      // (IEEE operations, so the device's per-particle values are identical)
      const double m = 1.0;
      const double alpha = m / P.tau;
      const double bvar = 2.0 * alpha * P.T_ * P.dt;
      const double qq = alpha * P.dt / 2.0 / m;
      const double num = 1.0 - qq;
      const double den = 1.0 + qq;
      P.g1_sq = std::sqrt(bvar);
      P.g1_ca = num / den;
      P.g1_cb = 1.0 / den;
    }
  }
  if (c0 >= (i64)0x7fffffff) return ctx->fail(DP_ERR_ARG, "too many cells (%lld)", (long long)c0);
  ctx->cmask = 0;
  for (const auto& P : reps) ctx->cmask |= P.collapsed_mask;
  // fixed-point scale: |values| < 2^30 units
  int qbits = (int)std::floor(30.0 - std::log2(rmax));
  if (qbits > 30) qbits = 30;
  ctx->qbits = qbits;
  const double qs = std::ldexp(1.0, qbits);
  for (auto& P : reps) {
    for (int d = 0; d < D; ++d) {
      P.Wq[d] = (i32)std::rint(P.w[d] * qs);
      P.Lq[d] = (i32)std::min(std::rint(P.L[d] * qs), 2147483647.0);
    }
    // emit screening: fixed-point distances differ from the exact ones by
    // less than 1.5 units per dimension (two roundings of offsets, one of the
    // cell width, one of the box length) -> a band of `margin` units around
    // the cutoff is decided by the FP64 test
    const double margin = 2.0 * std::sqrt((double)D) + 2.0;
    const double thr = P.nl_thresh;
    if (!(thr > 0.0)) {
      P.nlq_lo = 0;
      P.nlq_hi = 0;
    } else if (!std::isfinite(thr) || std::sqrt(thr) * qs + margin > 2.0e9 * std::sqrt((double)D)) {
      P.nlq_lo = 0;  // everything goes through the FP64 test (no fixed-point range)
      P.nlq_hi = ~(u64)0;
    } else {
      const double Rq = std::sqrt(thr) * qs;
      const double a = std::max(Rq - margin, 0.0), b = Rq + margin;
      P.nlq_lo = (u64)std::max(std::floor(a * a) - 1.0e4, 0.0);
      P.nlq_hi = (u64)(std::ceil(b * b) + 1.0e4);
    }
    // FP32 screen: D converted differences and their sum of squares carry a
    // relative error below 5 * 2^-24; 2^-20 of slack on each side covers it
    if (P.nlq_hi == 0) {
      P.nlq_lo_f = P.nlq_hi_f = -1.0f;  // nothing is listed
    } else if (P.nlq_lo == 0 && P.nlq_hi == ~(u64)0) {
      P.nlq_lo_f = -1.0f;  // everything takes the exact path
      P.nlq_hi_f = INFINITY;
    } else {
      float lf = (float)((double)P.nlq_lo * (1.0 - std::ldexp(1.0, -20)));
      float hf = (float)((double)P.nlq_hi * (1.0 + std::ldexp(1.0, -20)));
      P.nlq_lo_f = std::nextafter(lf, -INFINITY);
      P.nlq_hi_f = std::nextafter(hf, INFINITY);
    }
    // 16-bit screen: t_d = (a_d >> 16) - (b_d >> 16) is within 1 of the
    // displacement / 2^16, so with s16 = sum t_d^2 the exact r^2 / 2^32 lies in
    // [(sqrt(s16) - sqrt(D))^2 - D, (sqrt(s16) + sqrt(D))^2]; 2 units of slack
    // on each side cover the host's rounding
    if (P.nlq_hi == 0) {
      P.nlq16_lo = 0;  // nothing is listed
      P.nlq16_hi = -1;
    } else if (P.nlq_lo == 0 && P.nlq_hi == ~(u64)0) {
      P.nlq16_lo = 0;  // everything takes the exact path (s16 saturates at 2^31 - 1)
      P.nlq16_hi = 2147483647;
    } else {
      const double sD = std::sqrt((double)D);
      const double qlo = std::ldexp((double)P.nlq_lo, -32), qhi = std::ldexp((double)P.nlq_hi, -32);
      const double rl = std::sqrt(qlo) - sD;
      const double lo16 = rl > 0.0 ? std::floor(rl * rl) - 2.0 : 0.0;
      const double rh = std::sqrt(qhi + D) + sD;
      const double hi16 = std::ceil(rh * rh) + 2.0;
      P.nlq16_lo = (i32)std::max(lo16, 0.0);
      P.nlq16_hi = (i32)std::min(hi16, 2147483647.0);
    }
  }
  // entry format: the index of a neighbour in its tile's staged halo,
  // shifted left by esh; in mixed precision esh = log2 of the staged record's
  // bytes, so that an entry is the record's byte offset in the pair kernels'
  // LDS record area.  Mixed-precision rows are padded with the entry of a
  // sentinel record (staged after the halo, 2^31 fixed-point units from the
  // home box's centre along every dimension) where that is provably beyond
  // the cut of every home particle: along d no displacement from it is
  // smaller than b_d = 2^31 - (T_d Wq_d / 2 + skin 2^q + 2) (a home
  // particle's record lies within T_d Wq_d / 2 + skin 2^q of the box centre).
  // Not with a single-cell dimension (its minimum image would fold the
  // sentinel back).  The emit gives a tile this format when its halo fits
  // the pair kernels' LDS record area (at most 48 KB of records, so byte
  // offsets fit 16 bits), other tiles plain indices: 16-bit entries unless a
  // halo could exceed 2^16 records.
  (void)occ_max;
  int esh = 0, sent = 0;
  if (ctx->mode == DP_MIXED) {
    esh = staged_esh(D, ctx->mode, ctx->ntypes);
    sent = (ctx->cmask == 0) ? 1 : 0;
    for (int r = 0; r < ctx->R && sent; ++r) {
      const DpRep& P = reps[r];
      const double cut = ctx->cutoff[r], sk = std::max(ctx->skin[r], 0.0);
      if (!std::isfinite(cut)) {
        sent = 0;
        break;
      }
      long double s2 = 0.0L;
      for (int d = 0; d < D; ++d) {
        const long double b = 2147483648.0L - ((long double)P.T[d] * (long double)P.Wq[d] / 2.0L +
                                               (long double)sk * (long double)qs + 2.0L);
        if (!(b > 0.0L)) sent = 0;
        s2 += b * b;
      }
      const long double need = ((long double)cut + 2.0L * (long double)sk) * (long double)qs;
      if (!(s2 > need * need * 1.0001L + 1048576.0L)) sent = 0;
    }
  }
  const int efmt = esh | (sent << 8);
  int nlw = ctx->nlw_min;
  if (halo_occ_max * 4.0 + 64.0 >= 65536.0) nlw = 32;
  const int ob = nlw;
  // geometry of the built list: cells (a change needs a rebuild) and tiles /
  // entry width / fixed-point scale (a change needs a re-emit from xref)
  std::vector<i64> cells_sig, tiles_sig;
  for (const auto& P : reps) {
    cells_sig.push_back(P.c0);
    for (int d = 0; d < D; ++d) cells_sig.push_back(P.nc[d]);
    tiles_sig.push_back(P.t0);
    tiles_sig.push_back(P.nslice);
    for (int d = 0; d < D; ++d) tiles_sig.push_back(P.T[d]);
  }
  tiles_sig.push_back(nlw);
  tiles_sig.push_back(qbits);
  tiles_sig.push_back(efmt);
  const bool cells_changed = cells_sig != ctx->cells_sig;
  const bool tiles_changed = tiles_sig != ctx->tiles_sig;
  ctx->cells_sig = cells_sig;
  ctx->tiles_sig = tiles_sig;
  ctx->nlw = nlw;
  ctx->ob = ob;
  ctx->efmt = efmt;
  ctx->tpp = tpp;
  ctx->halo_occ = halo_occ_max;
  ctx->rec_bytes = rec_bytes;
  if (cells_changed) ctx->halo_seen = 0;  // halos measured on other cells
  size_lds(ctx);
  {
    // cells of a few particles: one emit lane per particle (the
    // candidate-table pass takes a wavefront per home cell and group)
    const int lanes = occ_max < 8.0 ? 1 : 0;
    if (lanes != ctx->emit_lanes) release_graph(ctx);  // captured emit launches carry the mode
    ctx->emit_lanes = lanes;
  }
  ctx->ntiles = t0;
  {
    // one workgroup per tile (measured: 209 -> 202 us for the 3-D LJ force
    // kernel at N = 2^20 against 4096 workgroups looping over tiles); the
    // kernels still loop when the grid is smaller than the tile count
    i64 g = std::min<i64>(t0, (i64)1 << 20);
    g = std::max<i64>(8, (g + 7) / 8 * 8);
    ctx->tile_grid = (int)g;
  }
  ctx->reps_h = reps;
  if (c0 != ctx->Ctot) {
    ctx->Ctot = c0;
    CHK(dalloc(ctx, ctx->ccount, (size_t)c0));
    HIPCHK(hipMemsetAsync(ctx->ccount, 0, (size_t)c0 * 4, ctx->st));
    CHK(dalloc(ctx, ctx->cstart, (size_t)c0 + 1));
    release_graph(ctx);
  }
  // tile tables: one fixed stride for all tiles
  {
    int ts = 0;
    for (const auto& P : reps) ts = std::max(ts, dp_tile_table_ints(P.nhalo, P.nhome));
    ts = (ts + 31) & ~31;
    const size_t bytes = (size_t)t0 * (size_t)ts * 4;
    if (ts != ctx->tstride || bytes != ctx->tiletab_bytes || !ctx->tiletab) {
      dfree(ctx->tiletab);
      ctx->tiletab_bytes = 0;
      if (hipMalloc(&ctx->tiletab, bytes) != hipSuccess) {
        ctx->tiletab = nullptr;
        return ctx->fail(DP_ERR_HIP, "cannot allocate tile tables of %zu bytes", bytes);
      }
      ctx->tiletab_bytes = bytes;
      ctx->tstride = ts;
      ctx->nl_valid = false;  // tables are written by the emit
      release_graph(ctx);
    }
  }
  if (t0 != ctx->todo_n || !ctx->todo) {
    CHK(dalloc(ctx, ctx->todo, (size_t)std::max<i64>(t0, 1)));
    ctx->todo_n = t0;
    release_graph(ctx);  // captured emits carry the list's address
  }
  {
    // batches: at most N / DP_BLOCK + R particle blocks of rebuilding replicas
    const i64 pn = ctx->N / DP_BLOCK + ctx->R + 1;
    if (pn != ctx->plist_n || !ctx->plist) {
      CHK(dalloc(ctx, ctx->plist, (size_t)pn));
      CHK(dalloc(ctx, ctx->rlist, (size_t)std::max(ctx->R, 1)));
      ctx->plist_n = pn;
      release_graph(ctx);  // captured rebuilds carry the lists' addresses
    }
  }
  // rows padded to a multiple of 8 entries (the pair kernels read 8 bytes at a time)
  const size_t nl_entries = (size_t)ctx->N * (size_t)((ctx->K + 7) & ~7);
  if (nl_entries >= ((size_t)1 << 32))  // the emit indexes the list with 32-bit entry offsets
    return ctx->fail(DP_ERR_ARG, "neighbour list of %zu entries exceeds 2^32 - 1 (N x max_number_of_neighbors)",
                     nl_entries);
  const size_t nl_bytes = nl_entries * (size_t)(nlw / 8);
  if (nl_bytes != ctx->nl_bytes || !ctx->nl) {
    dfree(ctx->nl);
    ctx->nl_bytes = 0;
    // + DP_NL_SLACK: the pair kernels load a row's next 16-byte group before
    // testing whether the row goes on (at most 8 x lanes-per-particle
    // entries past its end; the last row's loads end in this slack)
    if (hipMalloc(&ctx->nl, nl_bytes + DP_NL_SLACK) != hipSuccess) {
      ctx->nl = nullptr;
      return ctx->fail(DP_ERR_HIP, "cannot allocate neighbour list of %zu bytes", nl_bytes);
    }
    ctx->nl_bytes = nl_bytes;
    ctx->nl_valid = false;  // re-emitted below when a list exists
    release_graph(ctx);
  }
  if (cells_changed || tiles_changed) release_graph(ctx);
  if (cells_changed) ctx->nl_built = false;
  const bool reemit = ctx->nl_built && (tiles_changed || !ctx->nl_valid);
  HIPCHK(hipMemcpyAsync(ctx->reps_d, ctx->reps_h.data(), sizeof(DpRep) * ctx->R, hipMemcpyHostToDevice, ctx->st));
  ctx->reps_dirty = false;
  ctx->tab_dirty = true;  // guard bands depend on the fixed-point scale
  if (reemit) {
    // same cells and list, new tile plan: records and entries follow the plan
    CHK(launch_records(ctx));
    CHK(launch_reemit(ctx));
  }
  ctx->nl_valid = true;
  HIPCHK(hipStreamSynchronize(ctx->st));  // reps_h may change on the next call
  return DP_OK;
}

// Pair tables: exact FP64 cut thresholds and, for the mixed kernels, the
// guard band of the FP32 screening of fixed-point squared distances.  The
// fixed-point displacement differs from the exact one by at most 1.5 units
// per dimension (two rounded offsets and a rounded cell width); the FP32 sum
// of squares adds a few 2^-24 relative roundings.  Pairs inside the band are
// decided with the reference's FP64 arithmetic.
int refresh_tab(dp_ctx* ctx) {
  if (!ctx->tab_dirty) return DP_OK;
  if (!ctx->have_pot) return ctx->fail(DP_ERR_STATE, "no potential loaded (dp_load_potential)");
  const int nt = ctx->ntypes;
  const double qs = std::ldexp(1.0, ctx->qbits);
  const double e32 = std::ldexp(1.0, -24);
  for (DpPairTab& T : ctx->tabs_h) {
    T.qscale = 1.0 / qs;
    T.qscale2 = 1.0 / (qs * qs);
    T.qinv = qs;
    for (int a = 0; a < nt; ++a)
      for (int b = 0; b < nt; ++b) {
        const int k = a * nt + b;
        const double s = T.sigma[k], e = T.eps[k];
        T.inv_sigma2[k] = 1.0 / (s * s);
        T.eps_over_sigma[k] = e / s;
        T.inv_sigma2_qf[k] = (float)(T.inv_sigma2[k] / (qs * qs));
        T.eps_over_sigma_qf[k] = (float)(T.eps_over_sigma[k] / qs);
        T.eps_f[k] = (float)e;
        T.sigma_f[k] = (float)s;
        const double t = T.cut_thresh[k];
        if (!std::isfinite(t)) {
          T.guard_lo_q[k] = INFINITY;
          T.guard_hi_q[k] = INFINITY;
          continue;
        }
        if (t < 0) {
          T.guard_lo_q[k] = -1.0f;
          T.guard_hi_q[k] = -1.0f;
          continue;
        }
        const double tq = t * qs * qs;
        double g = 3.0 * std::sqrt((double)ctx->D) * std::sqrt(tq) + 2.25 * ctx->D + 8.0 * e32 * tq;
        g *= 4.0;  // safety factor
        float lo = (float)(tq - g), hi = (float)(tq + g);
        lo = std::nextafter(lo, -INFINITY);
        hi = std::nextafter(hi, INFINITY);
        if (lo < 0) lo = 0;
        T.guard_lo_q[k] = lo;
        T.guard_hi_q[k] = hi;
      }
  }
  // folded force coefficients: a_j = c_j * eps / sigma * sigma^(2 e_j) *
  // 2^((2 e_j - 1) q) per type pair (extended precision, then one rounding);
  // folding is off for a potential whose coefficients leave [2^-1000, 2^1000]
  for (size_t p = 0; p < ctx->tabs_h.size(); ++p) {
    DpPairTab& T = ctx->tabs_h[p];
    T.fold = 0;
    int n = 0, ex[4];
    double cf[4];
    if (ctx->mode != DP_MIXED || ctx->pair_prec != 64 || p >= ctx->pot_srcs.size() || !ctx->pot_evens[p] ||
        !parse_gsk(ctx->pot_srcs[p], n, ex, cf))
      continue;
    bool ok = true;
    for (int k = 0; k < nt * nt; ++k) {
      const long double sg = T.sigma[k], ep = T.eps[k];
      for (int j = 0; j < 4; ++j) {
        double a = 0.0;
        if (j < n) {
          const long double v = (long double)cf[j] * ep / sg * std::pow(sg * sg, (long double)ex[j]);
          a = (double)std::ldexp(v, (2 * ex[j] - 1) * ctx->qbits);
          if (!std::isfinite(a) || (a != 0.0 && std::fabs(std::log2(std::fabs(a))) > 1000.0)) ok = false;
        }
        T.gk[k * 4 + j] = a;
      }
    }
    T.fold = ok ? 1 : 0;
  }
  HIPCHK(hipMemcpyAsync(ctx->tab_d, ctx->tabs_h.data(), sizeof(DpPairTab) * ctx->tabs_h.size(),
                        hipMemcpyHostToDevice, ctx->st));
  HIPCHK(hipStreamSynchronize(ctx->st));
  ctx->tab_dirty = false;
  return DP_OK;
}

// Compile (or fetch) the pair module for the current plan.
int ensure_module(dp_ctx* ctx) {
  if (!ctx->have_pot) return ctx->fail(DP_ERR_STATE, "no potential loaded (dp_load_potential)");
  PairConfig c{ctx->D,  ctx->mode,      ctx->ntypes, 0, ctx->rs, ctx->tpp, ctx->nlw, ctx->ob, ctx->pair_prec,
               ctx->cmask, ctx->efmt & 31, ctx->efmt >> 8};
  // tile tables in LDS sized for the plan's largest tile (rounded up to 16 cells)
  int hmax = 1, omax = 1;
  for (const auto& P : ctx->reps_h) {
    hmax = std::max(hmax, (int)P.nhalo);
    omax = std::max(omax, (int)P.nhome);
  }
  c.hmax = std::min((hmax + 15) / 16 * 16, DP_HALO_MAX);
  c.omax = std::min((omax + 15) / 16 * 16, DP_HOME_MAX);
  std::vector<int> folds;
  for (const DpPairTab& T : ctx->tabs_h) folds.push_back(T.fold);
  std::string src = pair_source(c, ctx->pot_srcs, ctx->pot_evens, folds);
  if (ctx->mod && src == ctx->mod_key) return DP_OK;
  std::vector<char> code;
  std::string log;
  if (compile_cached(src, code, log) != DP_OK) return ctx->fail(DP_ERR_RTC, "hiprtc compile failed: %s", log.c_str());
  release_graph(ctx);
  if (ctx->mod) {
    (void)hipModuleUnload(ctx->mod);
    ctx->mod = nullptr;
  }
  HIPCHK(hipModuleLoadData(&ctx->mod, code.data()));
  HIPCHK(hipModuleGetFunction(&ctx->f_force, ctx->mod, "dp_tile_force"));
  HIPCHK(hipModuleGetFunction(&ctx->f_obs, ctx->mod, "dp_tile_obs"));
  HIPCHK(hipModuleGetFunction(&ctx->f_gjf, ctx->mod, "dp_tile_gjf"));
  HIPCHK(hipModuleGetFunction(&ctx->f_lf, ctx->mod, "dp_tile_lf"));
  HIPCHK(hipModuleGetFunction(&ctx->f_int_gjf, ctx->mod, "dp_integrate_gjf"));
  HIPCHK(hipModuleGetFunction(&ctx->f_int_lf, ctx->mod, "dp_integrate_lf"));
  {
    int sh = 0;
    if (hipFuncGetAttribute(&sh, HIP_FUNC_ATTRIBUTE_SHARED_SIZE_BYTES, ctx->f_int_gjf) == hipSuccess && sh >= 0)
      ctx->integ_static_lds = sh;
  }
  ctx->mod_key = src;
  return DP_OK;
}

int prepare(dp_ctx* ctx) {
  CHK(refresh_reps(ctx));
  CHK(refresh_tab(ctx));
  CHK(ensure_module(ctx));
  return DP_OK;
}

DpPairArgs pair_args(dp_ctx* ctx, int cur) {
  DpPairArgs a;
  std::memset(&a, 0, sizeof(a));
  a.x = ctx->x[cur];
  a.xq = ctx->xq[cur];
  a.nl = ctx->nl;
  a.nlcnt = ctx->nlcnt;
  a.type = ctx->type;
  a.cstart = ctx->cstart;
  a.reps = ctx->reps_d;
  a.tab = ctx->tab_d;
  a.N = ctx->N;
  a.ntiles = ctx->ntiles;
  a.K = ctx->K;
  a.R = ctx->R;
  a.cap = ctx->cap;
  a.tiletab = ctx->tiletab;
  a.tstride = ctx->tstride;
  a.x_out = ctx->x[1 - cur];
  a.xq_out = ctx->xq[1 - cur];
  a.v = ctx->v;
  a.beta = ctx->beta;
  a.img = ctx->img;
  a.mass = ctx->mass;
  a.xref = ctx->xref;
  a.cellpack = ctx->cellpack;
  a.orig = ctx->orig;
  a.rep = ctx->rep;
  a.xi = nullptr;
  a.seed = ctx->seed;  // (the kernels key the noise with each replica's DpRep::seed)
  a.ctl = ctx->ctl_d;
  a.blk_pos = -1;
  a.wrap = 1;
  a.rflag = ctx->rflag;
  a.unit_mass = ctx->unit_mass ? 1 : 0;
  return a;
}

// Tile kernels: grid-stride over tiles with dynamic LDS for the staged halo.
int launch_pair(dp_ctx* ctx, hipFunction_t f, DpPairArgs& a) {
  void* params[] = {&a};
  HIPCHK(hipModuleLaunchKernel(f, (unsigned)ctx->tile_grid, 1, 1, DP_TILE_THREADS, 1, 1,
                               (unsigned)ctx->lds_pair, ctx->st, params, nullptr));
  return DP_OK;
}

// Per-particle kernels (integrator with given forces): one workgroup per 256
// particles, at five workgroups per compute unit (the kernels' launch bound),
// or at four where the last round at five would be under a third full (too
// few workgroups left to keep HBM busy) and four per unit waste less: rounds
// x workgroups per unit, the time of a bandwidth-bound launch, is the smaller.
// At N = 2^20, 4096 workgroups make 4.0 rounds at four per unit against 3.2
// at five (measured 53.0 -> 49.8 us); at 2^21 the last round at five is 40 %
// full and four per unit measured 1 % slower.  Four per unit are imposed
// through the LDS request, more than a fifth of the unit's LDS.
int launch_flat(dp_ctx* ctx, hipFunction_t f, DpPairArgs& a) {
  const i64 nblk = (ctx->N + 255) / 256;
  const unsigned grid = (unsigned)nblk;
  const i64 slots = (i64)std::max(ctx->ncu, 1);
  const i64 t4 = (nblk + 4 * slots - 1) / (4 * slots) * 4, t5 = (nblk + 5 * slots - 1) / (5 * slots) * 5;
  const i64 last5 = nblk % (5 * slots);  // workgroups in the last round at five per unit (0: a full round)
  unsigned lds = 0;
  if (t4 < t5 && last5 != 0 && 3 * last5 < 5 * slots) {
    const int want = ctx->lds_cu / 5 + 1024 - ctx->integ_static_lds;  // > 1/5 of the unit in all
    if (want > 0 && ctx->integ_static_lds + want <= ctx->lds_cu / 4) lds = (unsigned)want;
  }
  void* params[] = {&a};
  HIPCHK(hipModuleLaunchKernel(f, grid, 1, 1, 256, 1, 1, lds, ctx->st, params, nullptr));
  return DP_OK;
}

// Neighbour-list emit over the current cells from xref; force = 1 runs it
// outside a rebuild (same list, new tile plan).
template <int D, typename E>
int launch_emit_te(dp_ctx* ctx, int force) {
  const double qinv = std::ldexp(1.0, ctx->qbits);
  // batches: only the rebuilding replicas' tiles, listed on the device, a
  // bounded grid looping over the list (a grid of one workgroup per tile
  // would spend most of a partial rebuild starting and retiring idle ones)
  const bool listed = !force && ctx->R > 1;
  int grid = ctx->tile_grid;
  if (listed) {
    k_emit_list<<<1, 256, 0, ctx->st>>>(ctx->reps_d, ctx->R, ctx->rflag + ctx->R, ctx->todo, ctx->ctl_d);
    grid = std::min(grid, 1024);
  }
  const i32* todo = listed ? ctx->todo : nullptr;
  k_emit_tile<D, E><<<grid, DP_TILE_THREADS, ctx->lds_emit, ctx->st>>>(
      ctx->xref, ctx->cellpack, qinv, ctx->cstart, ctx->reps_d, ctx->R, ctx->ntiles, ctx->K, ctx->cap_emit,
      (E*)ctx->nl, ctx->nlcnt, ctx->tiletab, ctx->tstride, ctx->ctl_d, force, ctx->emit_lanes, todo, ctx->efmt,
      ctx->rec_bytes ? (int)(48 * 1024 / ctx->rec_bytes) : ctx->cap);
  // (a bounded grid: the kernel usually finds no marked tile)
  k_emit_tile_far<D, E><<<std::min(grid, 1024), DP_TILE_THREADS, 0, ctx->st>>>(
      ctx->xref, ctx->cellpack, qinv, ctx->reps_d, ctx->ntiles, ctx->K, (E*)ctx->nl, ctx->nlcnt, ctx->tiletab,
      ctx->tstride, ctx->ctl_d, force, todo, ctx->efmt);
  HIPCHK(hipGetLastError());
  return DP_OK;
}

template <int D>
int launch_emit_t(dp_ctx* ctx, int force) {
  return ctx->nlw == 16 ? launch_emit_te<D, u16>(ctx, force) : launch_emit_te<D, u32>(ctx, force);
}

int launch_reemit(dp_ctx* ctx) {
  k_clear_stats<<<1, 256, 0, ctx->st>>>(ctx->ctl_d);
  switch (ctx->D) {
    case 1: return launch_emit_t<1>(ctx, 1);
    case 2: return launch_emit_t<2>(ctx, 1);
    case 3: return launch_emit_t<3>(ctx, 1);
    case 4: return launch_emit_t<4>(ctx, 1);
  }
  return ctx->fail(DP_ERR_ARG, "dimension %d unsupported", ctx->D);
}

template <int D>
int launch_rebuild_t(dp_ctx* ctx) {
  const int cur = ctx->cur;
  const i64 N = ctx->N;
  // one thread per particle: these kernels wait on memory, and a capped grid
  // (a few wavefronts per SIMD, each looping) leaves that latency exposed
  const int gN = grid_for(N, 1 << 20);
  hipStream_t st = ctx->st;
  // the cell counts are zero here: cleared at allocation and by the scan
  const i32* rb = ctx->rflag + ctx->R;  // this step's rebuild flag per replica
  // batches: only the rebuilding replicas' particles (a device-built list of
  // their particle blocks, a bounded grid) and cells (a scan per replica)
  const bool listed = ctx->R > 1;
  const int2* pl = listed ? ctx->plist : nullptr;
  const int gP = listed ? std::min(gN, 2048) : gN;
  if (listed)
    k_rebuild_list<<<1, 256, 0, st>>>(ctx->reps_d, ctx->R, rb, ctx->plist, ctx->rlist, ctx->ctl_d);
  k_bin<D><<<gP, DP_BLOCK, 0, st>>>(ctx->x[cur], ctx->rep, ctx->reps_d, N, ctx->cell_of, ctx->rank, ctx->ccount,
                                     ctx->ctl_d, rb, pl);
  if (listed) {
    k_scan_reps<<<std::min(ctx->R, 256), DP_SCAN_THREADS, 0, st>>>(ctx->ccount, ctx->cstart, ctx->reps_d, ctx->rlist,
                                                                   ctx->ctl_d);
  } else {
    // cell-count scan: one workgroup for few blocks of 16384 cells, three
    // passes over many workgroups otherwise (block sums in the scratch
    // buffer, free until the scatter)
    const i64 nb = (ctx->Ctot + DP_SCAN_PER * DP_SCAN_THREADS - 1) / (DP_SCAN_PER * DP_SCAN_THREADS);
    if (nb <= 4 || (size_t)nb * 4 > ctx->scratch_bytes) {
      k_scan_single<<<1, DP_SCAN_THREADS, 0, st>>>(ctx->ccount, ctx->Ctot, ctx->cstart, N, ctx->ctl_d);
    } else {
      i32* part = (i32*)ctx->scratch;
      k_scan_blocks<<<(unsigned)nb, DP_SCAN_THREADS, 0, st>>>(ctx->ccount, ctx->Ctot, part, ctx->ctl_d);
      k_scan_top<<<1, DP_SCAN_THREADS, 0, st>>>(part, nb, ctx->ctl_d);
      k_scan_apply<<<(unsigned)nb, DP_SCAN_THREADS, 0, st>>>(ctx->ccount, ctx->Ctot, ctx->cstart, part, N,
                                                            ctx->ctl_d);
    }
  }
  // the original ids in cell-grouped order go to the head of the scratch
  // buffer, which k_permute overwrites only after the in-cell sort
  i32* keys = (i32*)ctx->scratch;
  k_scatter<<<gP, DP_BLOCK, 0, st>>>(ctx->cell_of, ctx->rank, ctx->cstart, ctx->orig, N, ctx->tmp, keys, ctx->ctl_d,
                                     ctx->rep, rb, pl);
  k_cellsort<<<gP, DP_BLOCK, 0, st>>>(ctx->tmp, keys, ctx->cell_of, ctx->cstart, N, ctx->perm, ctx->ctl_d, ctx->rep,
                                      rb, pl);
  k_permute<D><<<gP, DP_BLOCK, 0, st>>>(ctx->perm, N, ctx->x[cur], ctx->v, ctx->beta, ctx->mass, ctx->img,
                                         ctx->type, ctx->orig, ctx->cell_of, ctx->scratch, ctx->ctl_d, ctx->rep, rb,
                                         pl);
  k_unpermute_store<D><<<gP, DP_BLOCK, 0, st>>>(N, ctx->x[cur], ctx->v, ctx->beta, ctx->mass, ctx->img, ctx->type,
                                                 ctx->orig, ctx->cell_of, ctx->scratch, ctx->xref, ctx->cellpack,
                                                 ctx->xq[cur], ctx->rs, std::ldexp(1.0, ctx->qbits), ctx->rep,
                                                 ctx->reps_d, ctx->ctl_d, rb, pl);
  CHK(launch_emit_t<D>(ctx, 0));
  return DP_OK;
}

int launch_rebuild(dp_ctx* ctx) {
  switch (ctx->D) {
    case 1: return launch_rebuild_t<1>(ctx);
    case 2: return launch_rebuild_t<2>(ctx);
    case 3: return launch_rebuild_t<3>(ctx);
    case 4: return launch_rebuild_t<4>(ctx);
  }
  return ctx->fail(DP_ERR_ARG, "dimension %d unsupported", ctx->D);
}

template <int D>
int launch_skin_check_t(dp_ctx* ctx) {
  // the flag of the last integrator pass refers to positions replaced since
  k_clear_need<<<1, 256, 0, ctx->st>>>(ctx->ctl_d, ctx->rflag, ctx->R);
  k_skin_check<D><<<grid_for(ctx->N), DP_BLOCK, 0, ctx->st>>>(ctx->x[ctx->cur], ctx->xref, ctx->rep, ctx->reps_d,
                                                              ctx->N, ctx->ctl_d, ctx->rflag);
  HIPCHK(hipGetLastError());
  return DP_OK;
}

int launch_skin_check(dp_ctx* ctx) {
  switch (ctx->D) {
    case 1: return launch_skin_check_t<1>(ctx);
    case 2: return launch_skin_check_t<2>(ctx);
    case 3: return launch_skin_check_t<3>(ctx);
    case 4: return launch_skin_check_t<4>(ctx);
  }
  return ctx->fail(DP_ERR_ARG, "dimension %d unsupported", ctx->D);
}

template <int D>
int launch_records_t(dp_ctx* ctx) {
  if (ctx->mode != DP_MIXED || !ctx->nl_built) return DP_OK;
  k_make_records<D><<<grid_for(ctx->N), DP_BLOCK, 0, ctx->st>>>(ctx->N, ctx->x[ctx->cur], ctx->type, ctx->cellpack,
                                                                ctx->xq[ctx->cur], ctx->rs, std::ldexp(1.0, ctx->qbits),
                                                                ctx->rep, ctx->reps_d);
  HIPCHK(hipGetLastError());
  return DP_OK;
}

int launch_records(dp_ctx* ctx) {
  switch (ctx->D) {
    case 1: return launch_records_t<1>(ctx);
    case 2: return launch_records_t<2>(ctx);
    case 3: return launch_records_t<3>(ctx);
    case 4: return launch_records_t<4>(ctx);
  }
  return ctx->fail(DP_ERR_ARG, "dimension %d unsupported", ctx->D);
}

// Fold the per-slot neighbour-list statistics into max_count / max_cell /
// max_halo of a host copy of the control block.
void reduce_ctl(DpCtl* c) {
  c->max_count = c->max_cell = c->max_halo = 0;
  for (int k = 0; k < DP_STAT_SLOTS; ++k) {
    const i32* sl = c->slot + k * DP_STAT_STRIDE;
    c->max_count = std::max(c->max_count, sl[DP_STAT_COUNT]);
    c->max_cell = std::max(c->max_cell, sl[DP_STAT_CELL]);
    c->max_halo = std::max(c->max_halo, sl[DP_STAT_HALO]);
  }
}

int launch_prologue(dp_ctx* ctx, int advance) {
  k_prologue<<<1, 256, 0, ctx->st>>>(ctx->ctl_d, ctx->rflag, ctx->R, ctx->nlupd_d, advance);
  HIPCHK(hipGetLastError());
  return DP_OK;
}

// Read the control block; report overflow / non-finite positions.
int check_ctl(dp_ctx* ctx) {
  HIPCHK(hipMemcpyAsync(ctx->ctl_h, ctx->ctl_d, sizeof(DpCtl), hipMemcpyDeviceToHost, ctx->st));
  HIPCHK(hipStreamSynchronize(ctx->st));
  reduce_ctl(ctx->ctl_h);
  if (ctx->ctl_h->max_halo > ctx->halo_seen) {
    // launches after this read-back size their LDS record areas for it
    ctx->halo_seen = ctx->ctl_h->max_halo;
    if (ctx->rec_bytes) size_lds(ctx);
  }
  ctx->last_cell_overflow = false;
  if (ctx->ctl_h->overflow || ctx->ctl_h->nonfinite || ctx->ctl_h->cell_overflow) {
    const bool ov = ctx->ctl_h->overflow != 0;
    const bool co = ctx->ctl_h->cell_overflow != 0;
    const int mc = ctx->ctl_h->max_count;
    // clear the sticky flags
    DpCtl c = *ctx->ctl_h;
    c.overflow = 0;
    c.nonfinite = 0;
    c.cell_overflow = 0;
    HIPCHK(hipMemcpy(ctx->ctl_d, &c, sizeof(DpCtl), hipMemcpyHostToDevice));
    if (ov)
      return ctx->fail(DP_ERR_OVERFLOW,
                       "Number of neighbors (%d) exceeds max_number_of_neighbors=%d; "
                       "increase max_number_of_neighbors",
                       mc, ctx->K);
    if (co) {
      // the list is incomplete; later builds use 32-bit entries
      ctx->last_cell_overflow = true;
      ctx->nlw_min = 32;
      ctx->reps_dirty = true;
      ctx->nl_built = false;
      release_graph(ctx);
      return ctx->fail(DP_ERR_OVERFLOW,
                       "a tile's halo holds %d particles, more than the %d that %d-bit neighbour entries address; "
                       "the next build uses 32-bit entries",
                       ctx->ctl_h->max_halo, dp_efmt_hlim(0), ctx->nlw);
    }
    return ctx->fail(DP_ERR_NONFINITE, "non-finite particle position encountered");
  }
  return DP_OK;
}

// Skin check from the current positions when the host changed them, and a
// forced rebuild if no list exists yet.
int pre_step_checks(dp_ctx* ctx) {
  if (!ctx->nl_built) {
    k_request_rebuild<<<1, 1, 0, ctx->st>>>(ctx->ctl_d);
    ctx->nl_built = true;  // the first prologue below builds it
    ctx->pos_dirty = false;
    return DP_OK;
  }
  if (ctx->pos_dirty) {
    CHK(launch_skin_check(ctx));
    ctx->pos_dirty = false;
  }
  return DP_OK;
}

int get_events(dp_ctx* ctx, DpEv3& ev) {
  if (!ctx->ev_free.empty()) {
    ev = ctx->ev_free.back();
    ctx->ev_free.pop_back();
    return DP_OK;
  }
  for (auto& e : ev.e) HIPCHK(hipEventCreate(&e));
  return DP_OK;
}

// The step kernels: the force kernel and the integrate kernel (or the fused
// kernel), as step `pos` of a step block (pos < 0: a single step).  With
// `ev`, e[0] / e[1] / e[2] are recorded before the force kernel, between the
// two and after the integrate kernel (e[1] == e[2] for the fused kernel).
int launch_step_kernels(dp_ctx* ctx, int integ, const double* xi_dev, const DpEv3* ev, int pos) {
  DpPairArgs a = pair_args(ctx, ctx->cur);
  a.xi = xi_dev;
  a.blk_pos = pos;
  if (ev) HIPCHK(hipEventRecordWithFlags(ev->e[0], ctx->st, ctx->capturing ? hipEventRecordExternal : 0));
  if (ctx->fused) {
    CHK(launch_pair(ctx, (integ == DP_INTEGRATOR_GJF) ? ctx->f_gjf : ctx->f_lf, a));
    if (ev) HIPCHK(hipEventRecordWithFlags(ev->e[1], ctx->st, ctx->capturing ? hipEventRecordExternal : 0));
  } else {
    a.F = ctx->F;
    CHK(launch_pair(ctx, ctx->f_force, a));
    if (ev) HIPCHK(hipEventRecordWithFlags(ev->e[1], ctx->st, ctx->capturing ? hipEventRecordExternal : 0));
    CHK(launch_flat(ctx, (integ == DP_INTEGRATOR_GJF) ? ctx->f_int_gjf : ctx->f_int_lf, a));
  }
  if (ev) HIPCHK(hipEventRecordWithFlags(ev->e[2], ctx->st, ctx->capturing ? hipEventRecordExternal : 0));
  ctx->cur = 1 - ctx->cur;
  return DP_OK;
}

// One step: control, gated rebuild, then the step kernels.
int launch_one_step(dp_ctx* ctx, int integ, const double* xi_dev, const DpEv3* ev) {
  CHK(launch_prologue(ctx, 1));
  CHK(launch_rebuild(ctx));
  return launch_step_kernels(ctx, integ, xi_dev, ev, -1);
}

int launch_one_step_timed(dp_ctx* ctx, int integ, const double* xi_dev) {
  if (!ctx->timing) return launch_one_step(ctx, integ, xi_dev, nullptr);
  DpEv3 ev;
  CHK(get_events(ctx, ev));
  CHK(launch_one_step(ctx, integ, xi_dev, &ev));
  ctx->ev_pending.push_back(ev);
  return DP_OK;
}

int resolve_timing(dp_ctx* ctx) {
  if (ctx->ev_pending.empty()) return DP_OK;
  HIPCHK(hipStreamSynchronize(ctx->st));
  for (auto& ev : ctx->ev_pending) {
    float ms = 0.f, m1 = 0.f;
    HIPCHK(hipEventElapsedTime(&ms, ev.e[0], ev.e[2]));
    HIPCHK(hipEventElapsedTime(&m1, ev.e[0], ev.e[1]));
    ctx->time_ms += ms;
    ctx->time_force_ms += m1;
    ctx->time_integ_ms += ms - m1;
    ctx->time_launches += 1;
    ctx->ev_free.push_back(ev);
  }
  ctx->ev_pending.clear();
  return DP_OK;
}

// Whether step blocks run the (gated) rebuild pipeline before every step,
// deciding per replica, instead of once at the block's start and stopping at
// the first skin trigger: the default for several replicas, where some
// replica's trigger fires on most steps.
int block_rebuilds_per_step(const dp_ctx* ctx) {
  return ctx->block_mode >= 0 ? ctx->block_mode : (ctx->R > 1 ? 1 : 0);
}

// A step block starting from buffer parity `par`: block control, the gated
// rebuild pipeline once, `graph_block` gated steps, block end.  Steps after
// the first one whose skin trigger fired are skipped; the host reads how many
// ran and launches the next block from the resulting parity.
int capture_graph(dp_ctx* ctx, int integ, int par, int timed) {
  hipGraphExec_t& gx = ctx->gexec[timed][par];
  if (gx) (void)hipGraphExecDestroy(gx);
  gx = nullptr;
  if (timed) {
    while ((int)ctx->blk_ev.size() < ctx->graph_block) {
      DpEv3 ev;
      for (auto& e : ev.e) HIPCHK(hipEventCreate(&e));
      ctx->blk_ev.push_back(ev);
    }
  }
  hipGraph_t g = nullptr;
  const int saved_cur = ctx->cur;
  ctx->cur = par;
  HIPCHK(hipStreamBeginCapture(ctx->st, hipStreamCaptureModeThreadLocal));
  ctx->capturing = true;
  const int perstep = block_rebuilds_per_step(ctx);
  k_block_begin<<<1, 256, 0, ctx->st>>>(ctx->ctl_d, ctx->nlupd_d, ctx->rflag, ctx->R, perstep);
  int rc = perstep ? DP_OK : launch_rebuild(ctx);
  for (int s = 0; s < ctx->graph_block && rc == DP_OK; ++s) {
    if (perstep) {
      // this step's rebuild decision per replica, then the gated pipeline
      k_step_begin<<<1, 256, 0, ctx->st>>>(ctx->ctl_d, ctx->nlupd_d, ctx->rflag, ctx->R, s);
      rc = launch_rebuild(ctx);
      if (rc != DP_OK) break;
    }
    rc = launch_step_kernels(ctx, integ, nullptr, timed ? &ctx->blk_ev[s] : nullptr, s);
  }
  k_block_end<<<1, 64, 0, ctx->st>>>(ctx->ctl_d, ctx->graph_block);
  ctx->capturing = false;
  hipError_t e = hipStreamEndCapture(ctx->st, &g);
  ctx->cur = saved_cur;
  if (rc != DP_OK) {
    if (g) (void)hipGraphDestroy(g);
    return rc;
  }
  if (e != hipSuccess) return ctx->fail(DP_ERR_HIP, "hipStreamEndCapture: %s", hipGetErrorString(e));
  e = hipGraphInstantiate(&gx, g, nullptr, nullptr, 0);
  (void)hipGraphDestroy(g);
  if (e != hipSuccess) {
    gx = nullptr;
    return ctx->fail(DP_ERR_HIP, "hipGraphInstantiate: %s", hipGetErrorString(e));
  }
  ctx->gexec_integ[timed][par] = integ;
  return DP_OK;
}

// Timed step blocks: add the event times of the `done` steps that ran;
// false if the runtime cannot time events recorded by a graph.
bool collect_block_times(dp_ctx* ctx, int done) {
  double t = 0.0, tf = 0.0;
  for (int p = 0; p < done; ++p) {
    float ms = 0.f, m1 = 0.f;
    const DpEv3& ev = ctx->blk_ev[p];
    if (hipEventElapsedTime(&ms, ev.e[0], ev.e[2]) != hipSuccess ||
        hipEventElapsedTime(&m1, ev.e[0], ev.e[1]) != hipSuccess) {
      (void)hipGetLastError();
      return false;
    }
    t += ms;
    tf += m1;
  }
  ctx->time_ms += t;
  ctx->time_force_ms += tf;
  ctx->time_integ_ms += t - tf;
  ctx->time_launches += done;
  return true;
}

}  // namespace

// ===========================================================================
// C-ABI
// ===========================================================================
DP_API int dp_version(void) { return 100; }


// Folded force terms for the compile checks: wherever the mixed kernels with
// FP64 pair terms can use them (the tables decide at run time).
std::vector<int> check_folds(const std::vector<std::string>& sv, int precision, int pair_precision) {
  std::vector<int> f(sv.size(), 0);
  int n, e[4];
  double cf[4];
  for (size_t k = 0; k < sv.size(); ++k)
    f[k] = (precision == DP_MIXED && !pair_precision && parse_gsk(sv[k], n, e, cf)) ? 1 : 0;
  return f;
}
int check_fold(const char* src, int precision, int pair_precision) {
  return check_folds(std::vector<std::string>{src}, precision, pair_precision)[0];
}

// Default plan used by the device-free compile checks (mixed precision: byte-
// offset entries padded with the sentinel's, as in the usual plan).
PairConfig default_config(int dim, int precision, int ntypes, int even, int pair_precision) {
  PairConfig c{dim, precision, ntypes, even, rs_for_dim(dim), 8, 16, 16,
               pair_precision ? 32 : 64, 0};
  c.esh = staged_esh(dim, precision, ntypes);
  c.sent = precision == DP_MIXED ? 1 : 0;
  return c;
}

DP_API int dp_compile_check(int dim, int precision, int ntypes, int even, const char* device_src, int pair_precision,
                            char* log, int64_t log_len, int64_t* code_size) {
  if (dim < 1 || dim > DP_DMAX || !device_src) return DP_ERR_ARG;
  if (ntypes < 1 || ntypes > DP_MAXTYPES) return DP_ERR_ARG;
  std::string src = pair_source(default_config(dim, precision, ntypes, even, pair_precision), device_src,
                                check_fold(device_src, precision, pair_precision));
  std::vector<char> code;
  std::string lg;
  int rc = compile_cached(src, code, lg);
  if (log && log_len > 0) {
    std::strncpy(log, lg.c_str(), (size_t)log_len - 1);
    log[log_len - 1] = '\0';
  }
  if (code_size) *code_size = (int64_t)code.size();
  return rc;
}

DP_API int dp_compile_check_tuned(int dim, int precision, int ntypes, int even, const char* device_src,
                                  int pair_precision, int threads_per_particle, int entry_bits,
                                  int collapsed_mask, char* log, int64_t log_len, int64_t* code_size) {
  if (dim < 1 || dim > DP_DMAX || !device_src) return DP_ERR_ARG;
  if (ntypes < 1 || ntypes > DP_MAXTYPES) return DP_ERR_ARG;
  const int t = threads_per_particle;
  if (t != 1 && t != 2 && t != 4 && t != 8 && t != 16) return DP_ERR_ARG;
  if (entry_bits != 16 && entry_bits != 32) return DP_ERR_ARG;
  PairConfig c = default_config(dim, precision, ntypes, even, pair_precision);
  if (collapsed_mask < 0 || collapsed_mask >= (1 << dim)) return DP_ERR_ARG;
  c.tpp = t;
  c.nlw = entry_bits;
  c.ob = entry_bits;
  c.cmask = collapsed_mask;
  std::string src = pair_source(c, device_src, check_fold(device_src, precision, pair_precision));
  std::vector<char> code;
  std::string lg;
  int rc = compile_cached(src, code, lg);
  if (log && log_len > 0) {
    std::strncpy(log, lg.c_str(), (size_t)log_len - 1);
    log[log_len - 1] = '\0';
  }
  if (code_size) *code_size = (int64_t)code.size();
  return rc;
}

DP_API int dp_compile_check_potentials(int dim, int precision, int ntypes, int npot, const char* const* device_srcs,
                                       const int* even, int pair_precision, char* log, int64_t log_len,
                                       int64_t* code_size) {
  if (dim < 1 || dim > DP_DMAX || !device_srcs || !even) return DP_ERR_ARG;
  if (ntypes < 1 || ntypes > DP_MAXTYPES || npot < 1 || npot > DP_MAXPOT) return DP_ERR_ARG;
  std::vector<std::string> sv;
  std::vector<int> ev;
  for (int p = 0; p < npot; ++p) {
    if (!device_srcs[p]) return DP_ERR_ARG;
    sv.emplace_back(device_srcs[p]);
    ev.push_back(even[p] ? 1 : 0);
  }
  std::string src = pair_source(default_config(dim, precision, ntypes, 0, pair_precision), sv, ev,
                                check_folds(sv, precision, pair_precision));
  std::vector<char> code;
  std::string lg;
  int rc = compile_cached(src, code, lg);
  if (log && log_len > 0) {
    std::strncpy(log, lg.c_str(), (size_t)log_len - 1);
    log[log_len - 1] = '\0';
  }
  if (code_size) *code_size = (int64_t)code.size();
  return rc;
}

DP_API int dp_compile_code(int dim, int precision, int ntypes, int even, const char* device_src, int pair_precision,
                           char* out, int64_t out_len, int64_t* code_size) {
  if (dim < 1 || dim > DP_DMAX || !device_src) return DP_ERR_ARG;
  if (ntypes < 1 || ntypes > DP_MAXTYPES) return DP_ERR_ARG;
  std::string src = pair_source(default_config(dim, precision, ntypes, even, pair_precision), device_src,
                                check_fold(device_src, precision, pair_precision));
  std::vector<char> code;
  std::string lg;
  int rc = compile_cached(src, code, lg);
  if (rc != DP_OK) return rc;
  if (code_size) *code_size = (int64_t)code.size();
  if (out && out_len >= (int64_t)code.size()) std::memcpy(out, code.data(), code.size());
  return DP_OK;
}

DP_API const char* dp_last_error(const dp_ctx* ctx) {
  return ctx ? ctx->err.c_str() : g_create_error.c_str();
}

DP_API int dp_create(dp_ctx** out, int device, int dim, int64_t n_total, int n_replicas, int precision) {
  if (!out) return DP_ERR_ARG;
  *out = nullptr;
  if (dim < 1 || dim > DP_DMAX) {
    g_create_error = "dimension must be in 1..8";
    return DP_ERR_ARG;
  }
  if (n_total < 1 || n_replicas < 1 || n_replicas > n_total) {
    g_create_error = "need n_total >= n_replicas >= 1";
    return DP_ERR_ARG;
  }
  if (precision != DP_FP64 && precision != DP_MIXED) {
    g_create_error = "precision must be DP_FP64 or DP_MIXED";
    return DP_ERR_ARG;
  }
  const int sb = 2 * dim;
  if (sb < 32 && (uint64_t)n_total > ((uint64_t)1 << (32 - sb))) {
    g_create_error = "too many particles for the neighbour-list entry format at this dimension";
    return DP_ERR_ARG;
  }
  int ndev = 0;
  hipError_t e = hipGetDeviceCount(&ndev);
  if (e != hipSuccess || ndev == 0) {
    g_create_error = std::string("no HIP device available: ") + hipGetErrorString(e);
    return DP_ERR_HIP;
  }
  if (device < 0 || device >= ndev) {
    g_create_error = "device index out of range";
    return DP_ERR_ARG;
  }
  dp_ctx* ctx = new dp_ctx();
  ctx->dev = device;
  ctx->D = dim;
  ctx->N = n_total;
  ctx->R = n_replicas;
  ctx->mode = precision;
  ctx->rs = rs_for_dim(dim);
  auto bail = [&](int rc) {
    g_create_error = ctx->err;
    dp_destroy(ctx);
    return rc;
  };
#define CCHK(expr)                 \
  do {                             \
    int rc_ = (expr);              \
    if (rc_ != DP_OK) return bail(rc_); \
  } while (0)
  if (hipSetDevice(device) != hipSuccess) return bail(ctx->fail(DP_ERR_HIP, "hipSetDevice failed"));
  if (hipStreamCreateWithFlags(&ctx->st, hipStreamNonBlocking) != hipSuccess)
    return bail(ctx->fail(DP_ERR_HIP, "hipStreamCreate failed"));
  {
    int ncu = 0, lds = 0;
    if (hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, device) == hipSuccess && ncu > 0)
      ctx->ncu = ncu;
    if (hipDeviceGetAttribute(&lds, hipDeviceAttributeMaxSharedMemoryPerMultiprocessor, device) == hipSuccess &&
        lds > 0)
      ctx->lds_cu = lds;
  }
  {
    // device address of the allocation-free entry point's scratch, once per device
    std::string serr;
    const int rc = scratch_address(device, &ctx->na_scratch, serr);
    if (rc != DP_OK) return bail(ctx->fail(rc, "%s", serr.c_str()));
  }
  const i64 N = n_total, D = dim;
  CCHK(dalloc(ctx, ctx->reps_d, (size_t)n_replicas));
  CCHK(dalloc(ctx, ctx->tab_d, DP_MAXPOT));
  CCHK(dalloc(ctx, ctx->ctl_d, 1));
  CCHK(dalloc(ctx, ctx->nlupd_d, (size_t)n_replicas));
  CCHK(dalloc(ctx, ctx->rflag, (size_t)2 * n_replicas));
  CCHK(dalloc(ctx, ctx->red, (size_t)n_replicas));
  CCHK(dalloc(ctx, ctx->x[0], (size_t)(N * D)));
  CCHK(dalloc(ctx, ctx->x[1], (size_t)(N * D)));
  if (precision == DP_MIXED) {
    CCHK(dalloc(ctx, ctx->xq[0], (size_t)(N * ctx->rs)));
    CCHK(dalloc(ctx, ctx->xq[1], (size_t)(N * ctx->rs)));
  }
  CCHK(dalloc(ctx, ctx->v, (size_t)(N * D)));
  CCHK(dalloc(ctx, ctx->beta, (size_t)(N * D)));
  CCHK(dalloc(ctx, ctx->mass, (size_t)N));
  CCHK(dalloc(ctx, ctx->xref, (size_t)(N * D)));
  CCHK(dalloc(ctx, ctx->F, (size_t)(N * D)));
  CCHK(dalloc(ctx, ctx->E, (size_t)N));
  CCHK(dalloc(ctx, ctx->W, (size_t)N));
  CCHK(dalloc(ctx, ctx->LAP, (size_t)N));
  CCHK(dalloc(ctx, ctx->xi, (size_t)(N * D)));
  CCHK(dalloc(ctx, ctx->img, (size_t)(N * D)));
  CCHK(dalloc(ctx, ctx->type, (size_t)N));
  CCHK(dalloc(ctx, ctx->orig, (size_t)N));
  CCHK(dalloc(ctx, ctx->rep, (size_t)N));
  CCHK(dalloc(ctx, ctx->cellpack, (size_t)N));
  CCHK(dalloc(ctx, ctx->cell_of, (size_t)N));
  CCHK(dalloc(ctx, ctx->rank, (size_t)N));
  CCHK(dalloc(ctx, ctx->tmp, (size_t)N));
  CCHK(dalloc(ctx, ctx->perm, (size_t)N));
  CCHK(dalloc(ctx, ctx->nlcnt, (size_t)N));
  ctx->scratch_bytes = (size_t)(N * (3 * D * 8 + 8 + D * 4 + 12));
  CCHK(dalloc(ctx, ctx->scratch, ctx->scratch_bytes));
  CCHK(dalloc(ctx, ctx->stage, (size_t)(N * D * 8)));
  if (hipHostMalloc((void**)&ctx->ctl_h, sizeof(DpCtl), hipHostMallocDefault) != hipSuccess)
    return bail(ctx->fail(DP_ERR_HIP, "hipHostMalloc failed"));
  // defaults: identity order, one replica covering everything, zero state
  {
    std::vector<i32> ident((size_t)N);
    for (i64 i = 0; i < N; ++i) ident[i] = (i32)i;
    if (hipMemcpy(ctx->orig, ident.data(), N * 4, hipMemcpyHostToDevice) != hipSuccess)
      return bail(ctx->fail(DP_ERR_HIP, "hipMemcpy failed"));
    std::vector<i32> zi((size_t)(N * D), 0);
    std::vector<double> zd((size_t)(N * D), 0.0);
    std::vector<double> ones((size_t)N, 1.0);
    bool ok = hipMemcpy(ctx->img, zi.data(), N * D * 4, hipMemcpyHostToDevice) == hipSuccess &&
              hipMemcpy(ctx->type, zi.data(), N * 4, hipMemcpyHostToDevice) == hipSuccess &&
              hipMemcpy(ctx->x[0], zd.data(), N * D * 8, hipMemcpyHostToDevice) == hipSuccess &&
              hipMemcpy(ctx->v, zd.data(), N * D * 8, hipMemcpyHostToDevice) == hipSuccess &&
              hipMemcpy(ctx->beta, zd.data(), N * D * 8, hipMemcpyHostToDevice) == hipSuccess &&
              hipMemcpy(ctx->xref, zd.data(), N * D * 8, hipMemcpyHostToDevice) == hipSuccess &&
              hipMemcpy(ctx->mass, ones.data(), N * 8, hipMemcpyHostToDevice) == hipSuccess &&
              hipMemcpy(ctx->nlcnt, zi.data(), N * 4, hipMemcpyHostToDevice) == hipSuccess;
    DpCtl c;
    std::memset(&c, 0, sizeof(c));
    ok = ok && hipMemcpy(ctx->ctl_d, &c, sizeof(c), hipMemcpyHostToDevice) == hipSuccess;
    std::vector<u64> z64((size_t)n_replicas, 0);
    ok = ok && hipMemcpy(ctx->nlupd_d, z64.data(), 8 * n_replicas, hipMemcpyHostToDevice) == hipSuccess;
    std::vector<i32> z32((size_t)2 * n_replicas, 0);
    ok = ok && hipMemcpy(ctx->rflag, z32.data(), 8 * n_replicas, hipMemcpyHostToDevice) == hipSuccess;
    if (!ok) return bail(ctx->fail(DP_ERR_HIP, "initialising device buffers failed"));
  }
  ctx->offsets = {0, n_total};
  {
    // every particle in replica 0 until dp_set_replicas says otherwise
    std::vector<i32> r0((size_t)N, 0);
    if (hipMemcpy(ctx->rep, r0.data(), N * 4, hipMemcpyHostToDevice) != hipSuccess)
      return bail(ctx->fail(DP_ERR_HIP, "hipMemcpy failed"));
  }
#undef CCHK
  *out = ctx;
  return DP_OK;
}

DP_API int dp_destroy(dp_ctx* ctx) {
  if (!ctx) return DP_OK;
  (void)hipSetDevice(ctx->dev);
  if (ctx->st) (void)hipStreamSynchronize(ctx->st);
  release_graph(ctx);
  for (auto& ev : ctx->blk_ev)
    for (auto& e : ev.e) (void)hipEventDestroy(e);
  for (auto& ev : ctx->ev_pending)
    for (auto& e : ev.e) (void)hipEventDestroy(e);
  for (auto& ev : ctx->ev_free)
    for (auto& e : ev.e) (void)hipEventDestroy(e);
  if (ctx->mod) (void)hipModuleUnload(ctx->mod);
  dfree(ctx->reps_d); dfree(ctx->tab_d); dfree(ctx->ctl_d); dfree(ctx->nlupd_d); dfree(ctx->rflag);
  dfree(ctx->red); dfree(ctx->x[0]); dfree(ctx->x[1]); dfree(ctx->xq[0]); dfree(ctx->xq[1]);
  dfree(ctx->v); dfree(ctx->beta); dfree(ctx->mass); dfree(ctx->xref); dfree(ctx->F); dfree(ctx->E);
  dfree(ctx->W); dfree(ctx->LAP); dfree(ctx->xi); dfree(ctx->img); dfree(ctx->type); dfree(ctx->orig);
  dfree(ctx->rep); dfree(ctx->cellpack); dfree(ctx->cell_of); dfree(ctx->rank); dfree(ctx->tmp);
  dfree(ctx->perm); dfree(ctx->nlcnt); dfree(ctx->ccount); dfree(ctx->cstart);
  dfree(ctx->nl); dfree(ctx->tiletab); dfree(ctx->todo); dfree(ctx->scratch); dfree(ctx->stage);
  dfree(ctx->plist); dfree(ctx->rlist);
  if (ctx->ctl_h) (void)hipHostFree(ctx->ctl_h);
  if (ctx->st) (void)hipStreamDestroy(ctx->st);
  delete ctx;
  return DP_OK;
}

DP_API int dp_set_replicas(dp_ctx* ctx, const int64_t* offsets, const double* box) {
  if (!ctx || !offsets || !box) return DP_ERR_ARG;
  (void)hipSetDevice(ctx->dev);
  const int R = ctx->R;
  if (offsets[0] != 0 || offsets[R] != ctx->N) return ctx->fail(DP_ERR_ARG, "offsets must run from 0 to N");
  for (int r = 0; r < R; ++r)
    if (offsets[r + 1] <= offsets[r]) return ctx->fail(DP_ERR_ARG, "every replica needs at least one particle");
  for (int k = 0; k < R * ctx->D; ++k)
    if (!(box[k] > 0.0) || !std::isfinite(box[k])) return ctx->fail(DP_ERR_ARG, "box lengths must be positive");
  const bool same_offsets = std::equal(ctx->offsets.begin(), ctx->offsets.end(), offsets) &&
                            (int)ctx->offsets.size() == R + 1;
  ctx->offsets.assign(offsets, offsets + R + 1);
  ctx->box.assign(box, box + (size_t)R * ctx->D);
  ctx->have_box = true;
  ctx->reps_dirty = true;
  ctx->nl_built = false;  // box changed: the list is rebuilt at the next use
  HIPCHK(hipStreamSynchronize(ctx->st));
  if (!same_offsets) {
    std::vector<i32> rv((size_t)ctx->N);
    for (int r = 0; r < R; ++r)
      for (i64 i = offsets[r]; i < offsets[r + 1]; ++i) rv[i] = r;
    HIPCHK(hipMemcpy(ctx->rep, rv.data(), ctx->N * 4, hipMemcpyHostToDevice));
  }
  release_graph(ctx);
  return DP_OK;
}

DP_API int dp_set_integrator(dp_ctx* ctx, const double* dt, const double* temperature, const double* damping_time) {
  if (!ctx || !dt || !temperature || !damping_time) return DP_ERR_ARG;
  const int R = ctx->R;
  ctx->dt.assign(dt, dt + R);
  ctx->T.assign(temperature, temperature + R);
  ctx->tau.assign(damping_time, damping_time + R);
  ctx->have_integ = true;
  ctx->reps_dirty = true;
  return DP_OK;
}

DP_API int dp_set_neighbor_params(dp_ctx* ctx, const double* cutoff, const double* skin, int max_neighbors,
                                  int order) {
  if (!ctx || !cutoff || !skin) return DP_ERR_ARG;
  (void)hipSetDevice(ctx->dev);
  if (max_neighbors < 1) return ctx->fail(DP_ERR_ARG, "max_number_of_neighbors must be >= 1");
  if (order != DP_ORDER_X_FASTEST && order != DP_ORDER_LAST_FASTEST) return ctx->fail(DP_ERR_ARG, "bad order");
  const int R = ctx->R;
  for (int r = 0; r < R; ++r)
    if (std::isnan(cutoff[r]) || std::isnan(skin[r])) return ctx->fail(DP_ERR_ARG, "cutoff/skin must not be NaN");
  ctx->cutoff.assign(cutoff, cutoff + R);
  ctx->skin.assign(skin, skin + R);
  ctx->order = order;
  if (max_neighbors != ctx->K) {
    ctx->K = max_neighbors;
    release_graph(ctx);
  }
  ctx->have_nl_params = true;
  ctx->reps_dirty = true;
  ctx->nl_built = false;
  return DP_OK;
}

namespace {
// Potentials k = 0..npot-1 (generated device functions srcs[k], even[k],
// r_cut[k], u(r_cut) u_rc[k], sigma / eps tables [k][ntypes][ntypes]);
// replica r uses potential rep_pot[r] (null: all use 0).
int load_potentials(dp_ctx* ctx, int npot, const char* const* srcs, const int* even, const double* r_cut,
                    const double* u_rc, int ntypes, const double* sigma, const double* eps, int pair_precision,
                    const int32_t* rep_pot) {
  (void)hipSetDevice(ctx->dev);
  if (npot < 1 || npot > DP_MAXPOT) return ctx->fail(DP_ERR_ARG, "number of potentials must be in 1..%d", DP_MAXPOT);
  if (ntypes < 1 || ntypes > DP_MAXTYPES) return ctx->fail(DP_ERR_ARG, "ntypes must be in 1..%d", DP_MAXTYPES);
  if (pair_precision != 0 && pair_precision != 1) return ctx->fail(DP_ERR_ARG, "pair_precision must be 0 or 1");
  const int ntt = ntypes * ntypes;
  for (int k = 0; k < npot * ntt; ++k)
    if (!(sigma[k] > 0.0)) return ctx->fail(DP_ERR_ARG, "sigma must be positive");
  for (int p = 0; p < npot; ++p)
    if (!srcs[p]) return ctx->fail(DP_ERR_ARG, "potential %d has no source", p);
  std::vector<i32> rp((size_t)ctx->R, 0);
  if (rep_pot)
    for (int r = 0; r < ctx->R; ++r) {
      if (rep_pot[r] < 0 || rep_pot[r] >= npot)
        return ctx->fail(DP_ERR_ARG, "replica %d: potential index %d out of range", r, rep_pot[r]);
      rp[r] = rep_pot[r];
    }
  std::vector<std::string> sv;
  std::vector<int> ev;
  for (int p = 0; p < npot; ++p) {
    sv.emplace_back(srcs[p]);
    ev.push_back(even[p] ? 1 : 0);
  }
  // compile now (with the default plan) so that errors surface here
  {
    PairConfig c = default_config(ctx->D, ctx->mode, ntypes, 0, pair_precision);
    std::string src = pair_source(c, sv, ev, check_folds(sv, ctx->mode, pair_precision));
    std::vector<char> code;
    std::string log;
    if (compile_cached(src, code, log) != DP_OK)
      return ctx->fail(DP_ERR_RTC, "hiprtc compile failed: %s", log.c_str());
  }
  ctx->pot_srcs = sv;
  ctx->pot_evens = ev;
  ctx->pair_prec = pair_precision ? 32 : 64;
  release_graph(ctx);
  if (ntypes != ctx->ntypes) ctx->reps_dirty = true;  // staged records carry the type with several types
  if (rp != ctx->rep_pot) ctx->reps_dirty = true;     // DpRep::pot
  ctx->rep_pot = rp;
  ctx->ntypes = ntypes;
  ctx->tabs_h.assign((size_t)npot, DpPairTab{});
  for (int p = 0; p < npot; ++p) {
    DpPairTab& T = ctx->tabs_h[p];
    std::memset(&T, 0, sizeof(T));
    for (int k = 0; k < ntt; ++k) {
      T.sigma[k] = sigma[(size_t)p * ntt + k];
      T.eps[k] = eps[(size_t)p * ntt + k];
      T.cut_thresh[k] = cut_threshold(T.sigma[k], r_cut[p]);
    }
    T.r_cut = r_cut[p];
    T.u_rc = u_rc[p];
  }
  ctx->have_pot = true;
  ctx->tab_dirty = true;
  return DP_OK;
}
}  // namespace

DP_API int dp_load_potential(dp_ctx* ctx, const char* device_src, int even, double r_cut, double u_rc, int ntypes,
                             const double* sigma, const double* eps, int pair_precision) {
  if (!ctx || !device_src || !sigma || !eps) return DP_ERR_ARG;
  const char* srcs[1] = {device_src};
  const int ev[1] = {even};
  return load_potentials(ctx, 1, srcs, ev, &r_cut, &u_rc, ntypes, sigma, eps, pair_precision, nullptr);
}

DP_API int dp_load_potentials(dp_ctx* ctx, int npot, const char* const* device_srcs, const int* even,
                              const double* r_cut, const double* u_rc, int ntypes, const double* sigma,
                              const double* eps, int pair_precision, const int32_t* rep_pot) {
  if (!ctx || !device_srcs || !even || !r_cut || !u_rc || !sigma || !eps || !rep_pot) return DP_ERR_ARG;
  return load_potentials(ctx, npot, device_srcs, even, r_cut, u_rc, ntypes, sigma, eps, pair_precision, rep_pot);
}

namespace {
int field_info(dp_ctx* ctx, int field, void** dev, size_t* elem, int* width) {
  switch (field) {
    case DP_FIELD_POSITIONS: *dev = ctx->x[ctx->cur]; *elem = 8; *width = ctx->D; return DP_OK;
    case DP_FIELD_VELOCITIES: *dev = ctx->v; *elem = 8; *width = ctx->D; return DP_OK;
    case DP_FIELD_BETAS: *dev = ctx->beta; *elem = 8; *width = ctx->D; return DP_OK;
    case DP_FIELD_IMAGES: *dev = ctx->img; *elem = 4; *width = ctx->D; return DP_OK;
    case DP_FIELD_MASSES: *dev = ctx->mass; *elem = 8; *width = 1; return DP_OK;
    case DP_FIELD_TYPES: *dev = ctx->type; *elem = 4; *width = 1; return DP_OK;
    case DP_FIELD_XREF: *dev = ctx->xref; *elem = 8; *width = ctx->D; return DP_OK;
  }
  return ctx->fail(DP_ERR_ARG, "unknown field %d", field);
}
}  // namespace

DP_API int dp_set_field(dp_ctx* ctx, int field, const void* src) {
  if (!ctx || !src) return DP_ERR_ARG;
  (void)hipSetDevice(ctx->dev);
  void* dev;
  size_t elem;
  int width;
  CHK(field_info(ctx, field, &dev, &elem, &width));
  const size_t bytes = (size_t)ctx->N * width * elem;
  HIPCHK(hipMemcpyAsync(ctx->stage, src, bytes, hipMemcpyDefault, ctx->st));
  const int g = grid_for(ctx->N);
  if (elem == 8)
    k_from_orig<double><<<g, DP_BLOCK, 0, ctx->st>>>((const double*)ctx->stage, (double*)dev, ctx->orig, ctx->N, width);
  else
    k_from_orig<i32><<<g, DP_BLOCK, 0, ctx->st>>>((const i32*)ctx->stage, (i32*)dev, ctx->orig, ctx->N, width);
  HIPCHK(hipGetLastError());
  if (field == DP_FIELD_TYPES && ctx->have_pot) {
    // type indices are used to address the sigma/epsilon tables
    std::vector<i32> t((size_t)ctx->N);
    HIPCHK(hipMemcpyAsync(t.data(), ctx->stage, ctx->N * 4, hipMemcpyDeviceToHost, ctx->st));
    HIPCHK(hipStreamSynchronize(ctx->st));
    for (i64 i = 0; i < ctx->N; ++i)
      if (t[i] < 0 || t[i] >= ctx->ntypes) return ctx->fail(DP_ERR_ARG, "particle type %d out of range", t[i]);
  }
  if (field == DP_FIELD_MASSES) {
    // unit masses let the integrators skip the mass array (a kernel argument:
    // captured step blocks are re-captured when it changes)
    std::vector<double> m((size_t)ctx->N);
    HIPCHK(hipMemcpyAsync(m.data(), ctx->stage, (size_t)ctx->N * 8, hipMemcpyDeviceToHost, ctx->st));
    HIPCHK(hipStreamSynchronize(ctx->st));
    bool unit = true;
    for (double v : m) unit = unit && (v == 1.0);
    if (unit != ctx->unit_mass) release_graph(ctx);
    ctx->unit_mass = unit;
  }
  if (field == DP_FIELD_POSITIONS || field == DP_FIELD_XREF) ctx->pos_dirty = true;
  if (field == DP_FIELD_POSITIONS || field == DP_FIELD_TYPES) CHK(launch_records(ctx));
  HIPCHK(hipStreamSynchronize(ctx->st));
  return DP_OK;
}

DP_API int dp_get_field(dp_ctx* ctx, int field, void* dst) {
  if (!ctx || !dst) return DP_ERR_ARG;
  (void)hipSetDevice(ctx->dev);
  void* dev;
  size_t elem;
  int width;
  CHK(field_info(ctx, field, &dev, &elem, &width));
  const size_t bytes = (size_t)ctx->N * width * elem;
  const int g = grid_for(ctx->N);
  if (elem == 8)
    k_to_orig<double><<<g, DP_BLOCK, 0, ctx->st>>>((const double*)dev, (double*)ctx->stage, ctx->orig, ctx->N, width);
  else
    k_to_orig<i32><<<g, DP_BLOCK, 0, ctx->st>>>((const i32*)dev, (i32*)ctx->stage, ctx->orig, ctx->N, width);
  HIPCHK(hipGetLastError());
  HIPCHK(hipMemcpyAsync(dst, ctx->stage, bytes, hipMemcpyDefault, ctx->st));
  HIPCHK(hipStreamSynchronize(ctx->st));
  return DP_OK;
}

DP_API int dp_build_neighbor_list(dp_ctx* ctx) {
  if (!ctx) return DP_ERR_ARG;
  (void)hipSetDevice(ctx->dev);
  int rc = DP_OK;
  for (int attempt = 0; attempt < 2; ++attempt) {
    CHK(refresh_reps(ctx));
    k_request_rebuild<<<1, 1, 0, ctx->st>>>(ctx->ctl_d);
    CHK(launch_prologue(ctx, 0));
    CHK(launch_rebuild(ctx));
    ctx->nl_built = true;
    ctx->pos_dirty = false;
    rc = check_ctl(ctx);
    // a halo beyond 16-bit entries: rebuilt once with 32-bit entries
    if (!(rc == DP_ERR_OVERFLOW && ctx->last_cell_overflow)) break;
  }
  return rc;
}

DP_API int dp_update_neighbor_list(dp_ctx* ctx, int* rebuilt) {
  if (!ctx) return DP_ERR_ARG;
  (void)hipSetDevice(ctx->dev);
  CHK(refresh_reps(ctx));
  if (!ctx->nl_built) {
    k_request_rebuild<<<1, 1, 0, ctx->st>>>(ctx->ctl_d);
    ctx->nl_built = true;
  } else {
    CHK(launch_skin_check(ctx));
  }
  ctx->pos_dirty = false;
  CHK(launch_prologue(ctx, 0));
  CHK(launch_rebuild(ctx));
  CHK(check_ctl(ctx));
  if (rebuilt) *rebuilt = ctx->ctl_h->rebuild;
  return DP_OK;
}

DP_API int dp_compute(dp_ctx* ctx, double* forces, double* energy, double* virial, double* laplacian) {
  if (!ctx) return DP_ERR_ARG;
  (void)hipSetDevice(ctx->dev);
  CHK(prepare(ctx));
  int rebuilt = 0;
  CHK(dp_update_neighbor_list(ctx, &rebuilt));
  DpPairArgs a = pair_args(ctx, ctx->cur);
  const bool obs = energy || virial || laplacian;
  a.F = ctx->F;
  if (obs) {
    a.E = energy ? ctx->E : nullptr;
    a.W = virial ? ctx->W : nullptr;
    a.LAP = laplacian ? ctx->LAP : nullptr;
    CHK(launch_pair(ctx, ctx->f_obs, a));
  } else {
    CHK(launch_pair(ctx, ctx->f_force, a));
  }
  const int g = grid_for(ctx->N);
  if (forces) {
    k_to_orig<double><<<g, DP_BLOCK, 0, ctx->st>>>(ctx->F, (double*)ctx->stage, ctx->orig, ctx->N, ctx->D);
    HIPCHK(hipMemcpyAsync(forces, ctx->stage, (size_t)ctx->N * ctx->D * 8, hipMemcpyDefault, ctx->st));
  }
  std::vector<double> tmp((size_t)ctx->R);
  if (energy) {
    k_reduce_rep<<<ctx->R, DP_BLOCK, 0, ctx->st>>>(ctx->E, ctx->reps_d, ctx->red);
    HIPCHK(hipMemcpyAsync(tmp.data(), ctx->red, 8 * ctx->R, hipMemcpyDeviceToHost, ctx->st));
    HIPCHK(hipStreamSynchronize(ctx->st));
    for (int r = 0; r < ctx->R; ++r) energy[r] = tmp[r] / 2;  // potential.py:100
  }
  if (virial) {
    k_reduce_rep<<<ctx->R, DP_BLOCK, 0, ctx->st>>>(ctx->W, ctx->reps_d, ctx->red);
    HIPCHK(hipMemcpyAsync(tmp.data(), ctx->red, 8 * ctx->R, hipMemcpyDeviceToHost, ctx->st));
    HIPCHK(hipStreamSynchronize(ctx->st));
    for (int r = 0; r < ctx->R; ++r) virial[r] = tmp[r] / 2 / ctx->D;  // each pair counted twice; potential.py:279
  }
  if (laplacian) {
    k_to_orig<double><<<g, DP_BLOCK, 0, ctx->st>>>(ctx->LAP, (double*)ctx->stage, ctx->orig, ctx->N, 1);
    HIPCHK(hipMemcpyAsync(laplacian, ctx->stage, (size_t)ctx->N * 8, hipMemcpyDefault, ctx->st));
  }
  HIPCHK(hipGetLastError());
  HIPCHK(hipStreamSynchronize(ctx->st));
  return DP_OK;
}

DP_API int dp_step(dp_ctx* ctx, int64_t nsteps, int integrator, const double* xi) {
  if (!ctx) return DP_ERR_ARG;
  (void)hipSetDevice(ctx->dev);
  if (integrator != DP_INTEGRATOR_GJF && integrator != DP_INTEGRATOR_LEAPFROG)
    return ctx->fail(DP_ERR_ARG, "unknown integrator %d", integrator);
  if (nsteps < 0) return ctx->fail(DP_ERR_ARG, "nsteps must be >= 0");
  if (nsteps == 0) return DP_OK;
  if (xi && nsteps != 1) return ctx->fail(DP_ERR_ARG, "injected noise requires nsteps == 1");
  if (!ctx->have_integ) return ctx->fail(DP_ERR_STATE, "dp_set_integrator has not been called");
  CHK(prepare(ctx));
  CHK(pre_step_checks(ctx));
  if (xi) {
    HIPCHK(hipMemcpyAsync(ctx->xi, xi, (size_t)ctx->N * ctx->D * 8, hipMemcpyDefault, ctx->st));
    CHK(launch_one_step_timed(ctx, integrator, ctx->xi));
    CHK(resolve_timing(ctx));
    return check_ctl(ctx);
  }
  int64_t left = nsteps;
  // timing brackets the step kernels with events: every step of one step
  // block in timing_every (the graph variant recording events), or every
  // single step where graphs cannot time events
  const int timed = ctx->timing ? 1 : 0;
  const bool use_graph = ctx->graph_block > 0 && (!timed || ctx->graph_events_ok) &&
                         nsteps >= (int64_t)ctx->graph_block / 2;
  if (use_graph) {
    if (ctx->graph_dirty) {
      release_graph(ctx);
      ctx->graph_dirty = false;
    }
    CHK(resolve_timing(ctx));
    while (left > 0) {
      const int par = ctx->cur;
      const int tb = (timed && ctx->timing_block++ % ctx->timing_every == 0) ? 1 : 0;
      if (!ctx->gexec[tb][par] || ctx->gexec_integ[tb][par] != integrator) {
        const int rc = capture_graph(ctx, integrator, par, tb);
        if (rc != DP_OK && tb) {
          // events inside captured graphs unavailable: time single steps
          ctx->graph_events_ok = false;
          (void)hipGetLastError();
          break;
        }
        CHK(rc);
      }
      k_block_left<<<1, 64, 0, ctx->st>>>(ctx->ctl_d, left);
      HIPCHK(hipGraphLaunch(ctx->gexec[tb][par], ctx->st));
      CHK(check_ctl(ctx));  // synchronises; the block's step count and errors
      const int done = ctx->ctl_h->blk_done;
      if (done <= 0 || done > left) return ctx->fail(DP_ERR_STATE, "step block completed %d steps", done);
      const bool times_ok = !tb || collect_block_times(ctx, done);
      left -= done;
      if (done & 1) ctx->cur = 1 - ctx->cur;
      if (!times_ok) {
        // events recorded by graphs cannot be timed here: time single steps
        ctx->graph_events_ok = false;
        break;
      }
    }
  }
  while (left > 0) {
    CHK(launch_one_step_timed(ctx, integrator, nullptr));
    --left;
    if (ctx->timing && ctx->ev_pending.size() >= 256) CHK(resolve_timing(ctx));
  }
  CHK(resolve_timing(ctx));
  return check_ctl(ctx);
}

DP_API int dp_integrate(dp_ctx* ctx, const double* forces, int integrator, const double* xi, int wrap) {
  if (!ctx || !forces) return DP_ERR_ARG;
  (void)hipSetDevice(ctx->dev);
  if (integrator != DP_INTEGRATOR_GJF && integrator != DP_INTEGRATOR_LEAPFROG)
    return ctx->fail(DP_ERR_ARG, "unknown integrator %d", integrator);
  if (!ctx->have_integ) return ctx->fail(DP_ERR_STATE, "dp_set_integrator has not been called");
  CHK(prepare(ctx));
  HIPCHK(hipMemcpyAsync(ctx->stage, forces, (size_t)ctx->N * ctx->D * 8, hipMemcpyDefault, ctx->st));
  k_from_orig<double><<<grid_for(ctx->N), DP_BLOCK, 0, ctx->st>>>((const double*)ctx->stage, ctx->F, ctx->orig,
                                                                  ctx->N, ctx->D);
  if (xi) HIPCHK(hipMemcpyAsync(ctx->xi, xi, (size_t)ctx->N * ctx->D * 8, hipMemcpyDefault, ctx->st));
  k_advance_step<<<1, 1, 0, ctx->st>>>(ctx->ctl_d);
  DpPairArgs a = pair_args(ctx, ctx->cur);
  a.F = ctx->F;
  a.xi = xi ? ctx->xi : nullptr;
  a.wrap = wrap ? 1 : 0;
  CHK(launch_flat(ctx, integrator == DP_INTEGRATOR_GJF ? ctx->f_int_gjf : ctx->f_int_lf, a));
  ctx->cur = 1 - ctx->cur;
  ctx->pos_dirty = true;  // the list's skin check must see the new positions
  CHK(launch_records(ctx));
  HIPCHK(hipStreamSynchronize(ctx->st));
  return DP_OK;
}

DP_API int dp_set_tuning(dp_ctx* ctx, int threads_per_particle, int tile_particles, int fused) {
  if (!ctx) return DP_ERR_ARG;
  if (threads_per_particle != 0 && threads_per_particle != 1 && threads_per_particle != 2 &&
      threads_per_particle != 4 && threads_per_particle != 8 && threads_per_particle != 16)
    return ctx->fail(DP_ERR_ARG, "threads_per_particle must be 0 (automatic), 1, 2, 4, 8 or 16");
  if (tile_particles < 0) return ctx->fail(DP_ERR_ARG, "tile_particles must be >= 0");
  ctx->tune_tpp = threads_per_particle;
  ctx->tune_tile = tile_particles;
  ctx->tune_fused = fused < 0 ? -1 : (fused ? 1 : 0);
  ctx->reps_dirty = true;
  release_graph(ctx);
  return DP_OK;
}

DP_API int dp_set_cell_scale(dp_ctx* ctx, double scale) {
  if (!ctx) return DP_ERR_ARG;
  if (!(scale >= 1.0) || !(scale <= 4.0)) return ctx->fail(DP_ERR_ARG, "cell scale must be in [1, 4]");
  if (scale != ctx->cell_scale) {
    ctx->cell_scale = scale;
    ctx->reps_dirty = true;
    ctx->nl_built = false;  // new cells: the list is rebuilt at the next use
    release_graph(ctx);
  }
  return DP_OK;
}

DP_API int dp_get_plan(dp_ctx* ctx, int32_t* out, int n) {
  if (!ctx || !out) return DP_ERR_ARG;
  (void)hipSetDevice(ctx->dev);
  CHK(refresh_reps(ctx));
  const int32_t v[16] = {ctx->tpp,  ctx->nlw, ctx->ob, ctx->qbits, ctx->cap, ctx->cap_emit,
                         (int32_t)ctx->ntiles, ctx->tile_grid, ctx->fused, ctx->reps_h[0].nhome,
                         ctx->reps_h[0].nhalo, (int32_t)ctx->Ctot, ctx->graph_block,
                         ctx->graph_events_ok ? 1 : 0, ctx->efmt & 31, ctx->efmt >> 8};
  for (int k = 0; k < n && k < 16; ++k) out[k] = v[k];
  return DP_OK;
}

DP_API int dp_set_seed(dp_ctx* ctx, uint64_t seed, uint64_t step_counter) {
  if (!ctx) return DP_ERR_ARG;
  (void)hipSetDevice(ctx->dev);
  ctx->seed = seed;
  ctx->rep_seeds.clear();
  ctx->reps_dirty = true;  // replica keys live in the replica table
  HIPCHK(hipStreamSynchronize(ctx->st));
  HIPCHK(hipMemcpy(ctx->ctl_h, ctx->ctl_d, sizeof(DpCtl), hipMemcpyDeviceToHost));
  ctx->ctl_h->step_counter = step_counter;
  HIPCHK(hipMemcpy(ctx->ctl_d, ctx->ctl_h, sizeof(DpCtl), hipMemcpyHostToDevice));
  release_graph(ctx);  // the seed is a kernel argument
  return DP_OK;
}

DP_API int dp_set_replica_seeds(dp_ctx* ctx, const uint64_t* seeds) {
  if (!ctx || !seeds) return DP_ERR_ARG;
  ctx->rep_seeds.assign(seeds, seeds + ctx->R);
  ctx->reps_dirty = true;
  release_graph(ctx);
  return DP_OK;
}

namespace {
template <int D>
int launch_export_t(dp_ctx* ctx, i32* out) {
  const int g = ctx->tile_grid;
  if (ctx->nlw == 16)
    k_nl_export<D, u16><<<g, DP_TILE_THREADS, 0, ctx->st>>>((const u16*)ctx->nl, ctx->nlcnt, ctx->orig, ctx->cstart,
                                                            ctx->reps_d, ctx->R, ctx->ntiles, ctx->K,
                                                            ctx->efmt & 31, ctx->tiletab, ctx->tstride, out);
  else
    k_nl_export<D, u32><<<g, DP_TILE_THREADS, 0, ctx->st>>>((const u32*)ctx->nl, ctx->nlcnt, ctx->orig, ctx->cstart,
                                                            ctx->reps_d, ctx->R, ctx->ntiles, ctx->K,
                                                            ctx->efmt & 31, ctx->tiletab, ctx->tstride, out);
  HIPCHK(hipGetLastError());
  return DP_OK;
}

int launch_export(dp_ctx* ctx, i32* out) {
  switch (ctx->D) {
    case 1: return launch_export_t<1>(ctx, out);
    case 2: return launch_export_t<2>(ctx, out);
    case 3: return launch_export_t<3>(ctx, out);
    case 4: return launch_export_t<4>(ctx, out);
  }
  return ctx->fail(DP_ERR_ARG, "dimension %d unsupported", ctx->D);
}
}  // namespace

DP_API int dp_get_neighbor_list(dp_ctx* ctx, int32_t* out) {
  if (!ctx || !out) return DP_ERR_ARG;
  (void)hipSetDevice(ctx->dev);
  if (!ctx->nl_built) CHK(dp_build_neighbor_list(ctx));
  const size_t bytes = (size_t)ctx->N * ctx->K * 4;
  i32* tmpbuf = nullptr;
  HIPCHK(hipMalloc((void**)&tmpbuf, bytes));
  CHK(launch_export(ctx, tmpbuf));
  hipError_t e = hipMemcpyAsync(out, tmpbuf, bytes, hipMemcpyDefault, ctx->st);
  hipError_t e2 = hipStreamSynchronize(ctx->st);
  (void)hipFree(tmpbuf);
  if (e != hipSuccess || e2 != hipSuccess) return ctx->fail(DP_ERR_HIP, "neighbour-list export failed");
  return DP_OK;
}

DP_API int dp_get_neighbor_counts(dp_ctx* ctx, int32_t* out) {
  if (!ctx || !out) return DP_ERR_ARG;
  (void)hipSetDevice(ctx->dev);
  if (!ctx->nl_built) CHK(dp_build_neighbor_list(ctx));
  k_to_orig<i32><<<grid_for(ctx->N), DP_BLOCK, 0, ctx->st>>>(ctx->nlcnt, (i32*)ctx->stage, ctx->orig, ctx->N, 1);
  HIPCHK(hipMemcpyAsync(out, ctx->stage, (size_t)ctx->N * 4, hipMemcpyDefault, ctx->st));
  HIPCHK(hipStreamSynchronize(ctx->st));
  return DP_OK;
}

namespace {
// Largest squared displacement since the last rebuild.  bm_dev: a device
// buffer of at least grid_for(N) (<= 512) doubles, or null (allocated here).
template <int D>
int rdf_max_disp(dp_ctx* ctx, double* disp_sq, double* bm_dev) {
  const int g = grid_for(ctx->N);
  double* bm = bm_dev;
  if (!bm && hipMalloc(&bm, (size_t)g * 8) != hipSuccess) return ctx->fail(DP_ERR_HIP, "cannot allocate %d doubles", g);
  k_rdf_disp<D><<<g, DP_BLOCK, 0, ctx->st>>>(ctx->x[ctx->cur], ctx->xref, ctx->reps_d, ctx->N, bm);
  double h[512];  // grid_for caps the grid at 512 workgroups
  int rc = DP_OK;
  if (hipGetLastError() != hipSuccess ||
      hipMemcpyAsync(h, bm, (size_t)g * 8, hipMemcpyDeviceToHost, ctx->st) != hipSuccess ||
      hipStreamSynchronize(ctx->st) != hipSuccess)
    rc = ctx->fail(DP_ERR_HIP, "displacement reduction failed");
  if (!bm_dev) (void)hipFree(bm);
  double m = 0.0;
  for (int k = 0; k < g && rc == DP_OK; ++k)
    if (!(h[k] <= m)) m = h[k];
  *disp_sq = m;
  return rc;
}

// Smallest double s with sqrt(s) >= e (correctly rounded, monotone sqrt):
// the squared-distance threshold of bin edge e.
double rdf_lower(double e) {
  if (!(e > 0.0)) return -INFINITY;  // every s >= 0 qualifies
  double s = e * e;
  while (s > 0.0 && std::sqrt(s) >= e) s = std::nextafter(s, -INFINITY);
  while (!(std::sqrt(s) >= e)) s = std::nextafter(s, INFINITY);
  return s;
}

// Largest double s with sqrt(s) <= e.
double rdf_upper(double e) {
  if (e < 0.0) return -INFINITY;  // no s >= 0 qualifies
  if (e == 0.0) return 0.0;
  double s = e * e;
  while (s < INFINITY && std::sqrt(s) <= e) s = std::nextafter(s, INFINITY);
  while (!(std::sqrt(s) <= e)) s = std::nextafter(s, -INFINITY);
  return s;
}

template <int D>
int launch_rdf_t(dp_ctx* ctx, const double* edges, int ne, unsigned long long* hist_d, const double* thr_d,
                 double u_last, double* bm_dev) {
  // the tiles of the last rebuild hold every pair up to the last edge when
  // edges[-1] + 2 * max displacement stays below every cell width
  const DpRep& P = ctx->reps_h[0];
  bool cells = ctx->nl_built;
  for (int d = 0; d < D; ++d) cells = cells && P.nc[d] >= 3;
  if (cells) {
    double dsq = 0.0;
    CHK(rdf_max_disp<D>(ctx, &dsq, bm_dev));
    const double reach = edges[ne - 1] + 2.0 * std::sqrt(dsq);
    for (int d = 0; d < D; ++d) cells = cells && reach <= P.w[d] * (1.0 - 1e-9);
  }
  // bins in chunks that fit the workgroup's LDS histogram
  const int nb = ne - 1;
  for (int c0 = 0; c0 < nb; c0 += DP_RDF_MAXBINS) {
    const int nbc = std::min(DP_RDF_MAXBINS, nb - c0);
    const int closed = (c0 + nbc == nb) ? 1 : 0;
    if (cells)
      k_rdf_cells<D><<<ctx->tile_grid, DP_TILE_THREADS, 0, ctx->st>>>(ctx->x[ctx->cur], ctx->reps_d, ctx->ntiles,
                                                                      ctx->tiletab, ctx->tstride, thr_d + c0, nbc + 1,
                                                                      closed, u_last, hist_d + c0);
    else
      k_rdf_all<D><<<2048, 256, 0, ctx->st>>>(ctx->x[ctx->cur], ctx->N, ctx->reps_d, thr_d + c0, nbc + 1, closed,
                                              u_last, hist_d + c0);
    HIPCHK(hipGetLastError());
  }
  ctx->rdf_path = cells ? 1 : 0;
  return DP_OK;
}
}  // namespace

DP_API int dp_rdf_counts(dp_ctx* ctx, const double* edges, int n_edges, uint64_t* counts) {
  if (!ctx || !edges || !counts) return DP_ERR_ARG;
  (void)hipSetDevice(ctx->dev);
  if (ctx->R != 1) return ctx->fail(DP_ERR_ARG, "dp_rdf_counts works on single-replica contexts");
  if (n_edges < 2) return ctx->fail(DP_ERR_ARG, "need at least 2 bin edges");
  for (int k = 1; k < n_edges; ++k)
    if (!(edges[k] >= edges[k - 1])) return ctx->fail(DP_ERR_ARG, "r_bins array must be sorted in ascending order.");
  CHK(refresh_reps(ctx));  // cells and tiles only: no potential needed
  std::vector<double> thr((size_t)n_edges);
  for (int k = 0; k < n_edges; ++k) thr[k] = rdf_lower(edges[k]);
  const double u_last = rdf_upper(edges[n_edges - 1]);
  double* edges_d = nullptr;
  unsigned long long* hist_d = nullptr;
  if (hipMalloc(&edges_d, (size_t)n_edges * 8) != hipSuccess ||
      hipMalloc(&hist_d, (size_t)(n_edges - 1) * 8) != hipSuccess) {
    if (edges_d) (void)hipFree(edges_d);
    return ctx->fail(DP_ERR_HIP, "cannot allocate the histogram");
  }
  int rc = DP_OK;
  if (hipMemcpyAsync(edges_d, thr.data(), (size_t)n_edges * 8, hipMemcpyHostToDevice, ctx->st) != hipSuccess ||
      hipMemsetAsync(hist_d, 0, (size_t)(n_edges - 1) * 8, ctx->st) != hipSuccess)
    rc = ctx->fail(DP_ERR_HIP, "histogram set-up failed");
  if (rc == DP_OK) {
    switch (ctx->D) {
      case 1: rc = launch_rdf_t<1>(ctx, edges, n_edges, hist_d, edges_d, u_last, nullptr); break;
      case 2: rc = launch_rdf_t<2>(ctx, edges, n_edges, hist_d, edges_d, u_last, nullptr); break;
      case 3: rc = launch_rdf_t<3>(ctx, edges, n_edges, hist_d, edges_d, u_last, nullptr); break;
      case 4: rc = launch_rdf_t<4>(ctx, edges, n_edges, hist_d, edges_d, u_last, nullptr); break;
    }
  }
  if (rc == DP_OK && (hipMemcpyAsync(counts, hist_d, (size_t)(n_edges - 1) * 8, hipMemcpyDeviceToHost, ctx->st) !=
                          hipSuccess ||
                      hipStreamSynchronize(ctx->st) != hipSuccess))
    rc = ctx->fail(DP_ERR_HIP, "histogram read-back failed");
  (void)hipFree(edges_d);
  (void)hipFree(hist_d);
  return rc;
}

DP_API int dp_rdf_counts_noalloc(dp_ctx* ctx, const double* edges, int n_edges, uint64_t* counts) {
  if (!ctx || !edges || !counts) return DP_ERR_ARG;
  (void)hipSetDevice(ctx->dev);
  if (ctx->R != 1) return ctx->fail(DP_ERR_ARG, "dp_rdf_counts_noalloc works on single-replica contexts");
  if (n_edges < 2) return ctx->fail(DP_ERR_ARG, "need at least 2 bin edges");
  if (n_edges > DP_NA_MAX_EDGES)
    return ctx->fail(DP_ERR_ARG, "dp_rdf_counts_noalloc takes at most %d bin edges (%d bins); got %d",
                     DP_NA_MAX_EDGES, DP_RDF_MAXBINS, n_edges);
  for (int k = 1; k < n_edges; ++k)
    if (!(edges[k] >= edges[k - 1])) return ctx->fail(DP_ERR_ARG, "r_bins array must be sorted in ascending order.");
  // a new tile plan may allocate: the plan must be current already
  if (ctx->reps_dirty || ctx->reps_h.empty())
    return ctx->fail(DP_ERR_STATE,
                     "the cell plan is not current: build the neighbour list (dp_build_neighbor_list) first");
  if (!ctx->na_scratch) return ctx->fail(DP_ERR_STATE, "no scratch address for device %d", ctx->dev);
  DpNoAllocLayout* sc = (DpNoAllocLayout*)ctx->na_scratch;  // device address
  double thr[DP_NA_MAX_EDGES];
  for (int k = 0; k < n_edges; ++k) thr[k] = rdf_lower(edges[k]);
  const double u_last = rdf_upper(edges[n_edges - 1]);
  const int nb = n_edges - 1;
  HIPCHK(hipMemcpyAsync(sc->thr, thr, (size_t)n_edges * 8, hipMemcpyHostToDevice, ctx->st));
  HIPCHK(hipStreamSynchronize(ctx->st));  // thr lives on this stack frame
  // the histogram keeps the last call's counts: cleared by a fill kernel
  k_fill_u64<<<(nb + 255) / 256, 256, 0, ctx->st>>>(sc->hist, nb);
  HIPCHK(hipGetLastError());
  int rc = DP_OK;
  switch (ctx->D) {
    case 1: rc = launch_rdf_t<1>(ctx, edges, n_edges, sc->hist, sc->thr, u_last, sc->block_max); break;
    case 2: rc = launch_rdf_t<2>(ctx, edges, n_edges, sc->hist, sc->thr, u_last, sc->block_max); break;
    case 3: rc = launch_rdf_t<3>(ctx, edges, n_edges, sc->hist, sc->thr, u_last, sc->block_max); break;
    case 4: rc = launch_rdf_t<4>(ctx, edges, n_edges, sc->hist, sc->thr, u_last, sc->block_max); break;
    default: rc = ctx->fail(DP_ERR_ARG, "dimension %d unsupported", ctx->D);
  }
  if (rc != DP_OK) return rc;
  HIPCHK(hipMemcpyAsync(counts, sc->hist, (size_t)nb * 8, hipMemcpyDeviceToHost, ctx->st));
  HIPCHK(hipStreamSynchronize(ctx->st));
  return DP_OK;
}

DP_API int dp_get_rdf_path(dp_ctx* ctx, int* cells) {
  if (!ctx || !cells) return DP_ERR_ARG;
  *cells = ctx->rdf_path;
  return DP_OK;
}

DP_API int dp_get_stats(dp_ctx* ctx, int64_t* nl_updates, int64_t* steps, int32_t* max_count) {
  if (!ctx) return DP_ERR_ARG;
  (void)hipSetDevice(ctx->dev);
  HIPCHK(hipStreamSynchronize(ctx->st));
  if (nl_updates) HIPCHK(hipMemcpy(nl_updates, ctx->nlupd_d, 8 * ctx->R, hipMemcpyDeviceToHost));
  HIPCHK(hipMemcpy(ctx->ctl_h, ctx->ctl_d, sizeof(DpCtl), hipMemcpyDeviceToHost));
  reduce_ctl(ctx->ctl_h);
  if (steps) *steps = (int64_t)ctx->ctl_h->step_counter;
  if (max_count) *max_count = ctx->ctl_h->max_count;
  return DP_OK;
}

DP_API int dp_set_graph_block(dp_ctx* ctx, int steps) {
  if (!ctx || steps < 0) return DP_ERR_ARG;
  if (steps > DP_BLK_MAX) return ctx->fail(DP_ERR_ARG, "step blocks hold at most %d steps", DP_BLK_MAX);
  ctx->graph_block = steps;
  release_graph(ctx);
  return DP_OK;
}

DP_API int dp_set_block_rebuild(dp_ctx* ctx, int mode) {
  if (!ctx) return DP_ERR_ARG;
  if (mode < -1 || mode > 1) return ctx->fail(DP_ERR_ARG, "block rebuild mode must be -1, 0 or 1");
  if (mode != ctx->block_mode) {
    ctx->block_mode = mode;
    release_graph(ctx);
  }
  return DP_OK;
}

DP_API int dp_prepare_step_graphs(dp_ctx* ctx, int integrator, int timed) {
  if (!ctx) return DP_ERR_ARG;
  (void)hipSetDevice(ctx->dev);
  if (integrator != DP_INTEGRATOR_GJF && integrator != DP_INTEGRATOR_LEAPFROG)
    return ctx->fail(DP_ERR_ARG, "unknown integrator %d", integrator);
  if (!ctx->have_integ) return ctx->fail(DP_ERR_STATE, "dp_set_integrator has not been called");
  if (ctx->graph_block <= 0) return DP_OK;
  const int t = timed ? 1 : 0;
  if (t && !ctx->graph_events_ok) return DP_OK;
  CHK(prepare(ctx));
  if (ctx->graph_dirty) {
    release_graph(ctx);
    ctx->graph_dirty = false;
  }
  for (int par = 0; par < 2; ++par) {
    if (ctx->gexec[t][par] && ctx->gexec_integ[t][par] == integrator) continue;
    const int rc = capture_graph(ctx, integrator, par, t);
    if (rc != DP_OK && t) {
      // events inside captured graphs unavailable: timed steps launch directly
      ctx->graph_events_ok = false;
      (void)hipGetLastError();
      return DP_OK;
    }
    CHK(rc);
  }
  return DP_OK;
}

DP_API int dp_set_kernel_timing(dp_ctx* ctx, int enable) {
  if (!ctx) return DP_ERR_ARG;
  (void)hipSetDevice(ctx->dev);
  CHK(resolve_timing(ctx));
  if (enable < 0) return ctx->fail(DP_ERR_ARG, "kernel timing must be 0 (off) or a block sampling interval >= 1");
  ctx->timing = enable != 0;
  ctx->timing_every = enable > 0 ? enable : 1;
  ctx->timing_block = 0;
  ctx->time_ms = 0.0;
  ctx->time_force_ms = 0.0;
  ctx->time_integ_ms = 0.0;
  ctx->time_launches = 0;
  return DP_OK;
}

DP_API int dp_get_kernel_time(dp_ctx* ctx, double* total_ms, int64_t* launches) {
  if (!ctx) return DP_ERR_ARG;
  (void)hipSetDevice(ctx->dev);
  CHK(resolve_timing(ctx));
  if (total_ms) *total_ms = ctx->time_ms;
  if (launches) *launches = ctx->time_launches;
  return DP_OK;
}

DP_API int dp_get_kernel_breakdown(dp_ctx* ctx, double* force_ms, double* integrate_ms) {
  if (!ctx) return DP_ERR_ARG;
  (void)hipSetDevice(ctx->dev);
  CHK(resolve_timing(ctx));
  if (force_ms) *force_ms = ctx->time_force_ms;
  if (integrate_ms) *integrate_ms = ctx->time_integ_ms;
  return DP_OK;
}

DP_API int dp_sync(dp_ctx* ctx) {
  if (!ctx) return DP_ERR_ARG;
  (void)hipSetDevice(ctx->dev);
  HIPCHK(hipStreamSynchronize(ctx->st));
  return DP_OK;
}

DP_API int dp_get_stream(dp_ctx* ctx, void** stream) {
  if (!ctx || !stream) return DP_ERR_ARG;
  *stream = (void*)ctx->st;
  return DP_OK;
}
