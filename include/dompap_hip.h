/* dompap_hip.h -- C-ABI of libdompap_hip.so, the MI355X (gfx950) engine for
 * dompap's per-step hot path: neighbour-list build, pair force / energy /
 * virial loop, and the G-JF Langevin / leap-frog update.
 *
 * The reference (dompap, pure Python + Numba) has no FFI; its operator API is
 * the Simulation method surface and the kernel free functions its tests
 * import.  Each entry point below names the reference interface it replaces
 * (file:line in the dompap repository).  The Python host in dompap_amd/ binds
 * these with ctypes (dompap_amd/_lib.py declares every entry point).
 *
 * Conventions
 *  - Every function returns an int status: DP_OK (0) or a negative DP_ERR_*;
 *    dp_last_error() returns the message.  No C++ exception crosses the ABI.
 *  - Per-particle arrays cross the ABI in the reference's layout and order:
 *    row n is particle n, [N][D] row-major (positions, velocities, betas,
 *    image_positions, forces), [N] for masses / types.  Pointers may be host
 *    memory or device memory (e.g. a torch tensor's data_ptr()); the library
 *    keeps its own device copy in cell-sorted order.
 *  - Several independent systems ("replicas") may share one context: replica
 *    r owns particles [offsets[r], offsets[r+1]).
 *  - One HIP stream per context; calls that return data synchronise.  A
 *    context is not thread-safe.
 */
#ifndef DOMPAP_HIP_H
#define DOMPAP_HIP_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define DP_OK 0
#define DP_ERR_ARG (-1)       /* bad argument (reference: ValueError) */
#define DP_ERR_HIP (-2)       /* HIP runtime error */
#define DP_ERR_RTC (-3)       /* hiprtc compile / module load error */
#define DP_ERR_OVERFLOW (-4)  /* a neighbour-list row exceeded max_number_of_neighbors */
#define DP_ERR_STATE (-5)     /* call out of order (e.g. no potential loaded) */
#define DP_ERR_NONFINITE (-6) /* a non-finite position was seen */

#define DP_FP64 0   /* pure FP64: matches the FP64 reference to 1e-10 relative */
#define DP_MIXED 1  /* int32 fixed-point coordinates relative to the cell centres (exact pair
                     * displacements); pair terms in FP64 by default (FP32 with pair_precision = 1
                     * in dp_load_potential); FP64 force accumulation, state, integration and
                     * energy / virial reductions */

#define DP_FIELD_POSITIONS 0  /* f64 [N][D]  Simulation.positions */
#define DP_FIELD_VELOCITIES 1 /* f64 [N][D]  Simulation.velocities */
#define DP_FIELD_BETAS 2      /* f64 [N][D]  Simulation.betas */
#define DP_FIELD_IMAGES 3     /* i32 [N][D]  Simulation.image_positions */
#define DP_FIELD_MASSES 4     /* f64 [N]     Simulation.masses[:, 0] */
#define DP_FIELD_TYPES 5      /* i32 [N]     type index into the sigma/eps tables */
#define DP_FIELD_XREF 6       /* f64 [N][D]  Simulation.positions_neighbour_list */

#define DP_INTEGRATOR_GJF 1      /* integrator.py:17-40 make_one_step */
#define DP_INTEGRATOR_LEAPFROG 2 /* integrator.py:5-14 make_one_step_leap_frog */

#define DP_ORDER_X_FASTEST 0    /* neighbor_list.py:150-157 stencil order (3-D cell list) */
#define DP_ORDER_LAST_FASTEST 1 /* neighbor_list.py:82-86 stencil order (any-D cell list) */

typedef struct dp_ctx dp_ctx;

/* Library version (major*10000 + minor*100 + patch). */
int dp_version(void);

/* Create a context for n_total particles in n_replicas systems of dimension
 * dim (1..4) on HIP device `device`.  Replaces the state storage of the
 * Simulation dataclass (dompap/Simulation.py:101-147). */
int dp_create(dp_ctx** out, int device, int dim, int64_t n_total, int n_replicas, int precision);
int dp_destroy(dp_ctx* ctx);
/* Message of the last failed call on ctx (or of the last failed dp_create
 * when ctx is NULL). */
const char* dp_last_error(const dp_ctx* ctx);

/* offsets[R+1] particle offsets, box[R*D] box lengths.
 * Simulation.box_vectors (Simulation.py:104). */
int dp_set_replicas(dp_ctx* ctx, const int64_t* offsets, const double* box);

/* Per-replica time step, target temperature, damping time (tau = +inf gives
 * NVE).  Simulation.set_integrator (Simulation.py:739-759). */
int dp_set_integrator(dp_ctx* ctx, const double* dt, const double* temperature, const double* damping_time);

/* Per-replica neighbour-list cutoff (max sigma * r_cut + skin,
 * Simulation.py:346-356) and skin; row capacity K
 * (max_number_of_neighbors); stencil order DP_ORDER_*.
 * Simulation.set_neighbor_list parameters (Simulation.py:307-345). */
int dp_set_neighbor_params(dp_ctx* ctx, const double* cutoff, const double* skin, int max_neighbors, int order);

/* Compile (hiprtc, gfx950) the pair kernels for a SymPy-generated potential.
 * device_src defines dp_pot_u/f/c (and dp_pot_gs/us when even != 0);
 * sigma/eps are ntypes*ntypes tables; u_rc = u(r_cut) (energy shift).
 * pair_precision (DP_MIXED only): 0 = in-range pair terms in FP64 from the
 * exact fixed-point displacement, 1 = in FP32; forces accumulate in FP64.
 * Replaces make_pair_potential (potential.py:43-66) and the sigma/epsilon
 * callables of set_pair_potential_parameters (Simulation.py:469-525). */
int dp_load_potential(dp_ctx* ctx, const char* device_src, int even, double r_cut, double u_rc, int ntypes,
                      const double* sigma, const double* eps, int pair_precision);

/* Several potentials in one context (batched replicas of different
 * potentials, stepped in the same launches): potential k has generated
 * device functions device_srcs[k], even[k], r_cut[k], u(r_cut) u_rc[k] and
 * sigma / eps tables [k][ntypes][ntypes]; replica r uses potential
 * rep_pot[r].  At most 8 potentials.  No counterpart in the reference (one
 * Simulation, one potential: potential.py:43-66, Simulation.py:403-525 per
 * replica). */
int dp_load_potentials(dp_ctx* ctx, int npot, const char* const* device_srcs, const int* even, const double* r_cut,
                       const double* u_rc, int ntypes, const double* sigma, const double* eps, int pair_precision,
                       const int32_t* rep_pot);

/* Upload / download one per-particle field in the reference layout.
 * Simulation attribute assignment / access (Simulation.py:101-147). */
int dp_set_field(dp_ctx* ctx, int field, const void* src);
int dp_get_field(dp_ctx* ctx, int field, void* dst);

/* Unconditional rebuild: Simulation.set_neighbor_list (Simulation.py:307-381)
 * with the cell-list pipeline; neighbour sets equal those of
 * get_neighbor_list_double_loop (neighbor_list.py:7-39). */
int dp_build_neighbor_list(dp_ctx* ctx);

/* Skin check then rebuild if old: Simulation.update_neighbor_list
 * (Simulation.py:383-401) with neighbor_list_is_old (neighbor_list.py:183-201).
 * *rebuilt (nullable) receives 1 if a rebuild happened. */
int dp_update_neighbor_list(dp_ctx* ctx, int* rebuilt);

/* Observables after the reference's update_neighbor_list():
 *  forces[N][D]  Simulation.get_forces -> _get_forces (potential.py:163-188)
 *  energy[R]     Simulation.get_potential_energy -> _get_total_energy (potential.py:69-100)
 *  virial[R]     Simulation.get_virial -> _get_virial_double_loop (potential.py:253-279)
 *  laplacian[N]  Simulation.get_laplacian -> _get_laplacian (potential.py:327-364)
 * Any output pointer may be NULL. */
int dp_compute(dp_ctx* ctx, double* forces, double* energy, double* virial, double* laplacian);

/* nsteps steps of Simulation.step (Simulation.py:681-704, G-JF) or
 * Simulation.step_leap_frog (:706-724): skin check, rebuild when old, forces,
 * integrator update, wrap (positions.py:90-98).  xi (nullable, only with
 * nsteps == 1) supplies the standard normals [N][D] of integrator.py:30;
 * otherwise they come from Philox4x32-10 keyed by (seed, step, particle). */
int dp_step(dp_ctx* ctx, int64_t nsteps, int integrator, const double* xi);

/* One integrator update with caller-supplied forces[N][D] (no pair loop):
 * make_one_step (integrator.py:17-40) / make_one_step_leap_frog
 * (integrator.py:5-14), followed by the wrap of positions.py:90-98 when
 * wrap != 0 (the reference's free functions do not wrap).  xi as in
 * dp_step. */
int dp_integrate(dp_ctx* ctx, const double* forces, int integrator, const double* xi, int wrap);

/* Kernel tuning (no counterpart in the reference's C surface; the GPU
 * autotuner replacing dompap/tools/autotune.py:4-178 drives it): lanes per particle
 * (0 = automatic, else 1/2/4/8/16), home particles per tile (0 = automatic),
 * fused force+integrate kernel (1), force kernel then integrate kernel (0) or
 * automatic (-1). */
int dp_set_tuning(dp_ctx* ctx, int threads_per_particle, int tile_particles, int fused);
/* Cells at least scale x cutoff wide (1 <= scale <= 4; default 1): larger
 * cells, fewer of them, larger tiles.  A tuning knob for the GPU autotuner
 * (dompap/tools/autotune.py:4-178 has no counterpart); lists and forces do
 * not depend on it beyond summation order.  The list is rebuilt at the next
 * use. */
int dp_set_cell_scale(dp_ctx* ctx, double scale);
/* Current plan: tpp, entry bits, offset bits, fixed-point bits, LDS record
 * capacity (pair, emit), tiles, tile grid, fused, home / halo cells per tile
 * of replica 0, cells, steps per step block, whether timed step blocks record
 * their events inside the graph, entry shift (an entry is a staged-halo index
 * shifted left by it), whether rows are padded with a sentinel record's entry
 * (first n values). */
int dp_get_plan(dp_ctx* ctx, int32_t* out, int n);

/* RNG key and step counter. */
int dp_set_seed(dp_ctx* ctx, uint64_t seed, uint64_t step_counter);
/* Noise key per replica (overrides dp_set_seed's key; dp_set_seed resets to
 * one key).  No counterpart in the reference: with it, replica r of a batch
 * draws the same numbers as a single-replica context keyed with seeds[r]
 * (numbers are keyed by the particle's index inside its replica). */
int dp_set_replica_seeds(dp_ctx* ctx, const uint64_t* seeds);

/* Neighbour list in the reference layout int32 [N][K], -1 padded
 * (Simulation.neighbor_list); rows in stencil order, members in ascending id.
 * (On the device the list is kept as 16- or 32-bit entries over cell-sorted
 * particles: the index of the neighbour in its tile's staged halo, shifted
 * left so that in mixed precision it is the byte offset of the neighbour's
 * staged record.) */
int dp_get_neighbor_list(dp_ctx* ctx, int32_t* out);
/* Row lengths int32 [N]: get_number_of_neighbors (neighbor_list.py:204-207). */
int dp_get_neighbor_counts(dp_ctx* ctx, int32_t* out);

/* Rebuild counts per replica (int64 [R]), device step counter, longest row. */
int dp_get_stats(dp_ctx* ctx, int64_t* nl_updates, int64_t* steps, int32_t* max_count);

/* Steps per captured hipGraph block used by dp_step (0 disables graphs). */
int dp_set_graph_block(dp_ctx* ctx, int steps);

/* How step blocks rebuild lists: 0 = the rebuild pipeline once at the
 * block's start, the block ending at the first step whose skin trigger fired
 * (one replica: fewest launches); 1 = the pipeline (gated) before every step,
 * only the replicas whose own trigger fired rebuilding, blocks never ending
 * early (many replicas); -1 = 1 iff the context holds several replicas
 * (default).  Replaces the per-step update_neighbor_list() call of
 * Simulation.step (Simulation.py:693) inside captured blocks. */
int dp_set_block_rebuild(dp_ctx* ctx, int mode);

/* Capture (without running) the step-block graphs dp_step replays from either
 * buffer parity, for `integrator` and the timed (events recorded inside the
 * graph) or untimed variant, so that a later dp_step spends no time on
 * capture and instantiation.  No counterpart in the reference (its run loop,
 * Simulation.py:726-737, has no graphs). */
int dp_prepare_step_graphs(dp_ctx* ctx, int integrator, int timed);

/* Kernel timing: enable = 0 turns it off; enable = k >= 1 brackets the step
 * kernels with HIP events on the context stream, recorded by the captured
 * step-block graph in the first step block and every k-th one after it (k = 1:
 * every step; a larger k samples whole blocks and keeps the events' own cost
 * off most steps), or around every step where graphs cannot time events.
 * dp_get_kernel_time returns the summed milliseconds and the number of
 * launches timed since timing was enabled. */
int dp_set_kernel_timing(dp_ctx* ctx, int enable);
int dp_get_kernel_time(dp_ctx* ctx, double* total_ms, int64_t* launches);

/* Compile the pair kernels for a potential without a device (hiprtc,
 * gfx950): build-time check of generated code.  log receives the compiler
 * log (truncated to log_len), code_size the code-object size.  Part of
 * make_pair_potential's replacement (potential.py:43-66). */
int dp_compile_check(int dim, int precision, int ntypes, int even, const char* device_src, int pair_precision,
                     char* log, int64_t log_len, int64_t* code_size);
/* Same check for an explicit kernel shape: lanes per particle (1, 2, 4, 8,
 * 16), neighbour-entry width (16 or 32 bits) and the mask of dimensions
 * holding a single cell, the variants dp_set_tuning and the tile planner can
 * select. */
int dp_compile_check_tuned(int dim, int precision, int ntypes, int even, const char* device_src,
                           int pair_precision, int threads_per_particle, int entry_bits, int collapsed_mask,
                           char* log, int64_t log_len, int64_t* code_size);
/* Same check for the kernels of several potentials in one context (as
 * dp_load_potentials compiles them). */
int dp_compile_check_potentials(int dim, int precision, int ntypes, int npot, const char* const* device_srcs,
                                const int* even, int pair_precision, char* log, int64_t log_len, int64_t* code_size);
/* Same compile; copies the gfx950 code object into out when out_len is large
 * enough (code_size receives the size). */
int dp_compile_code(int dim, int precision, int ntypes, int even, const char* device_src, int pair_precision,
                    char* out, int64_t out_len, int64_t* code_size);

/* Radial distribution function counts of a single-replica context
 * (positions.py:58-87, Simulation.py:858-881): counts[i] = number of ordered
 * pairs (n, m), n == m included, with edges[i] <= r_nm < edges[i+1] (last
 * bin closed), r_nm the minimum-image distance as NumPy computes it.  The
 * caller normalises by the shell volumes. */
int dp_rdf_counts(dp_ctx* ctx, const double* edges, int n_edges, uint64_t* counts);

/* dp_rdf_counts without any device or host allocation: no counterpart in the
 * reference.  For callers that cannot allocate (code under HIP graph capture,
 * tight serving loops, harnesses that forbid hipMalloc).  The thresholds,
 * histogram and displacement maxima live in a fixed device array of the
 * library (one per GPU, ~69 KB, kept until the process ends); at most 4097
 * edges (4096 bins), else DP_ERR_ARG.  The cell plan must be current (a
 * neighbour list built since the last change of box or list parameters),
 * else DP_ERR_STATE.  Calls that overlap on one GPU share the array and
 * race.  Results equal dp_rdf_counts's. */
int dp_rdf_counts_noalloc(dp_ctx* ctx, const double* edges, int n_edges, uint64_t* counts);

/* Which kernel the last dp_rdf_counts used: 1 = cell tiles of the current
 * list (last edge + 2 * max displacement < cell width), 0 = all pairs.  No
 * counterpart in the reference (its RDF is always all pairs). */
int dp_get_rdf_path(dp_ctx* ctx, int* cells);

/* Split of the timed step-kernel time: force (or fused) kernel and integrate
 * kernel, each summed over the timed steps (no counterpart in the reference). */
int dp_get_kernel_breakdown(dp_ctx* ctx, double* force_ms, double* integrate_ms);

/* Block until all work on the context stream is done. */
int dp_sync(dp_ctx* ctx);

/* The context's HIP stream (hipStream_t) for callers that order their own
 * work against it. */
int dp_get_stream(dp_ctx* ctx, void** stream);

#ifdef __cplusplus
}
#endif
#endif /* DOMPAP_HIP_H */
