/* dp_oracle.c -- C/OpenMP FP64 restatement of dompap's per-step hot path.
 *
 * TEST INFRASTRUCTURE: the oracle and the CPU baseline.  Only tests/,
 * __graft_entry__.smoke() and bench.py's cpu_baseline leg load it; the
 * product (dompap_amd) never does.
 *
 * Follows (dompap repository paths):
 *   neighbor_list.py:7-39     dpo_nl_double_loop (full list, ascending m,
 *                             strict sqrt(sum d^2) < cutoff, "<-L/2" first)
 *   neighbor_list.py:183-201  dpo_max_displacement2 / the skin test
 *   potential.py:43-66        closed-form u, f=-u', u'' for the potentials
 *                             the benchmarks use; u shifted by u(r_cut),
 *                             all three zero where x > r_cut
 *   potential.py:163-188      dpo_pair (forces)
 *   potential.py:69-100       dpo_pair (energy, halved)
 *   potential.py:253-279      dpo_pair (virial over the list, halved, /D)
 *   potential.py:327-364      dpo_pair (Laplacian)
 *   integrator.py:5-40        dpo_step (G-JF / leap-frog), numpy's
 *                             left-to-right operation order, no FMA
 *   positions.py:90-98        wrap inside dpo_step
 * dpo_nl_cells is a cell-list search producing the same rows as
 * dpo_nl_double_loop (ascending m); it substitutes for the O(N^2) loop where
 * that is infeasible (N >~ 1e5).
 * Compile with -ffp-contract=off (no fused multiply-add), as Numba does not
 * contract.
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#define DPO_LJ 0          /* 4*(r**-12 - r**-6) */
#define DPO_HARM2 1       /* (1-r)**2 */
#define DPO_HARM_HALF 2   /* "Harmonic repulsive": 0.5*(1-r)**2, f = 1-r, u'' = 0 (potential.py:10-28) */
#define DPO_IPL 3         /* r**-n, n = param */

typedef struct {
  int id;
  double n;      /* IPL exponent */
  double r_cut;
  double u_rc;
} dpo_pot;

static double ipow(double x, int n) {
  double r = 1.0;
  for (int k = 0; k < n; ++k) r = r * x;
  return r;
}

static double pot_u(const dpo_pot* p, double x) {
  switch (p->id) {
    case DPO_LJ: return 4.0 * (1.0 / ipow(x, 12) - 1.0 / ipow(x, 6));
    case DPO_HARM2: return (1.0 - x) * (1.0 - x);
    case DPO_HARM_HALF: return x < 1.0 ? 0.5 * (1.0 - x) * (1.0 - x) : 0.0;
    case DPO_IPL: return pow(x, -p->n);
  }
  return 0.0;
}
static double pot_f(const dpo_pot* p, double x) {
  switch (p->id) {
    case DPO_LJ: return 48.0 / ipow(x, 13) - 24.0 / ipow(x, 7);
    case DPO_HARM2: return 2.0 - 2.0 * x;
    case DPO_HARM_HALF: return x < 1.0 ? 1.0 - x : 0.0;
    case DPO_IPL: return p->n * pow(x, -p->n - 1.0);
  }
  return 0.0;
}
static double pot_c(const dpo_pot* p, double x) {
  switch (p->id) {
    case DPO_LJ: return 624.0 / ipow(x, 14) - 168.0 / ipow(x, 8);
    case DPO_HARM2: return 2.0;
    case DPO_HARM_HALF: return 0.0;
    case DPO_IPL: return p->n * (p->n + 1.0) * pow(x, -p->n - 2.0);
  }
  return 0.0;
}

/* u(r_cut) shift (potential.py:57-59); none for the hard-coded potential */
double dpo_shift(int id, double n, double r_cut) {
  dpo_pot p = {id, n, r_cut, 0.0};
  if (id == DPO_HARM_HALF || isinf(r_cut)) return (id == DPO_HARM2 && isinf(r_cut)) ? INFINITY : 0.0;
  return pot_u(&p, r_cut);
}

static inline double minimage_lt(double d, double L) {
  if (d < -L / 2) return d + L;
  if (d > L / 2) return d - L;
  return d;
}
static inline double minimage_gt(double d, double L) {
  if (d > L / 2) return d - L;
  if (d < -L / 2) return d + L;
  return d;
}

int dpo_num_threads(void) {
#ifdef _OPENMP
  return omp_get_max_threads();
#else
  return 1;
#endif
}

void dpo_set_threads(int n) {
#ifdef _OPENMP
  if (n > 0) omp_set_num_threads(n);
#else
  (void)n;
#endif
}

/* neighbor_list.py:7-39; returns the longest row (rows beyond K are not
 * written; the caller treats count > K as overflow). */
int dpo_nl_double_loop(int64_t N, int D, const double* x, const double* box, double cutoff, int K, int32_t* nl,
                       int32_t* counts) {
  int maxc = 0;
#pragma omp parallel for schedule(dynamic, 64) reduction(max : maxc)
  for (int64_t n = 0; n < N; ++n) {
    int c = 0;
    for (int64_t m = 0; m < N; ++m) {
      if (n == m) continue;
      double s = 0.0;
      for (int d = 0; d < D; ++d) {
        double t = minimage_lt(x[n * D + d] - x[m * D + d], box[d]);
        s = (d == 0) ? t * t : s + t * t;
      }
      if (sqrt(s) < cutoff) {
        if (c < K) nl[n * K + c] = (int32_t)m;
        ++c;
      }
    }
    for (int k = c; k < K; ++k) nl[n * K + k] = -1;
    counts[n] = c;
    if (c > maxc) maxc = c;
  }
  return maxc;
}

static int cmp_i32(const void* a, const void* b) {
  int32_t x = *(const int32_t*)a, y = *(const int32_t*)b;
  return (x > y) - (x < y);
}

/* Same rows as dpo_nl_double_loop from a cell search (cells >= cutoff wide;
 * a dimension with fewer than 3 cells is searched whole). */
int dpo_nl_cells(int64_t N, int D, const double* x, const double* box, double cutoff, int K, int32_t* nl,
                 int32_t* counts) {
  int nc[8];
  int64_t ncell = 1;
  for (int d = 0; d < D; ++d) {
    double q = floor(box[d] / cutoff);
    if (!(q >= 3.0)) q = 1.0;
    if (q > 1024) q = 1024;
    nc[d] = (int)q;
    ncell *= nc[d];
  }
  int64_t* cell = (int64_t*)malloc(sizeof(int64_t) * N);
  int64_t* start = (int64_t*)calloc((size_t)ncell + 1, sizeof(int64_t));
  int64_t* order = (int64_t*)malloc(sizeof(int64_t) * N);
  for (int64_t i = 0; i < N; ++i) {
    int64_t lin = 0, stride = 1;
    for (int d = 0; d < D; ++d) {
      double xd = x[i * D + d] - floor(x[i * D + d] / box[d]) * box[d];
      int c = (int)floor(xd * nc[d] / box[d]);
      if (c < 0) c = 0;
      if (c >= nc[d]) c = nc[d] - 1;
      lin += c * stride;
      stride *= nc[d];
    }
    cell[i] = lin;
    start[lin + 1]++;
  }
  for (int64_t c = 0; c < ncell; ++c) start[c + 1] += start[c];
  {
    int64_t* fill = (int64_t*)malloc(sizeof(int64_t) * (size_t)(ncell + 1));
    memcpy(fill, start, sizeof(int64_t) * (size_t)(ncell + 1));
    for (int64_t i = 0; i < N; ++i) order[fill[cell[i]]++] = i;
    free(fill);
  }
  int n3 = 1;
  for (int d = 0; d < D; ++d) n3 *= 3;
  int maxc = 0;
#pragma omp parallel for schedule(dynamic, 64) reduction(max : maxc)
  for (int64_t n = 0; n < N; ++n) {
    int cc[8];
    int64_t lin = cell[n];
    for (int d = 0; d < D; ++d) { cc[d] = (int)(lin % nc[d]); lin /= nc[d]; }
    int32_t* row = nl + n * K;
    int c = 0;
    for (int st = 0; st < n3; ++st) {
      int rem = st, skip = 0;
      int64_t l2 = 0, stride = 1;
      for (int d = 0; d < D; ++d) {
        int s = rem % 3 - 1;
        rem /= 3;
        if (nc[d] == 1 && s != 0) skip = 1;
        int q = cc[d] + s;
        if (q < 0) q += nc[d];
        else if (q >= nc[d]) q -= nc[d];
        l2 += q * stride;
        stride *= nc[d];
      }
      if (skip) continue;
      for (int64_t k = start[l2]; k < start[l2 + 1]; ++k) {
        int64_t m = order[k];
        if (m == n) continue;
        double s = 0.0;
        for (int d = 0; d < D; ++d) {
          double t = minimage_lt(x[n * D + d] - x[m * D + d], box[d]);
          s = (d == 0) ? t * t : s + t * t;
        }
        if (sqrt(s) < cutoff) {
          if (c < K) row[c] = (int32_t)m;
          ++c;
        }
      }
    }
    int w = c < K ? c : K;
    qsort(row, (size_t)w, sizeof(int32_t), cmp_i32);
    for (int k = w; k < K; ++k) row[k] = -1;
    counts[n] = c;
    if (c > maxc) maxc = c;
  }
  free(cell);
  free(start);
  free(order);
  return maxc;
}

/* neighbor_list.py:183-201: max over particles of |minimage(x - x_ref)|^2. */
double dpo_max_displacement2(int64_t N, int D, const double* x, const double* xref, const double* box) {
  double mx = 0.0;
#pragma omp parallel for reduction(max : mx)
  for (int64_t n = 0; n < N; ++n) {
    double s = 0.0;
    for (int d = 0; d < D; ++d) {
      double t = minimage_gt(x[n * D + d] - xref[n * D + d], box[d]);
      s = (d == 0) ? t * t : s + t * t;
    }
    if (s > mx) mx = s;
  }
  return mx;
}

/* Pair loop over a row-major list nl[N][K] (rows -1 terminated).
 * sigma/eps are per-type tables [T][T] indexed by types (NULL types = all 0).
 * Outputs (each nullable): F[N][D], e_pp[N] (per-particle eps*u sums),
 * w_pp[N] (eps*f*r sums), lap[N]. */
void dpo_pair(int64_t N, int D, const double* x, const double* box, const int32_t* nl, int K, int pot_id,
              double pot_n, double r_cut, int T, const int32_t* types, const double* sigma, const double* eps,
              double* F, double* e_pp, double* w_pp, double* lap) {
  dpo_pot p = {pot_id, pot_n, r_cut, dpo_shift(pot_id, pot_n, r_cut)};
#pragma omp parallel for schedule(dynamic, 256)
  for (int64_t n = 0; n < N; ++n) {
    double fn[8] = {0, 0, 0, 0, 0, 0, 0, 0};
    double e = 0.0, w = 0.0, l = 0.0;
    const int tn = types ? types[n] : 0;
    for (int k = 0; k < K; ++k) {
      const int32_t m = nl[n * K + k];
      if (m == -1) break;
      double dd[8];
      double s = 0.0;
      for (int d = 0; d < D; ++d) {
        dd[d] = minimage_gt(x[n * D + d] - x[(int64_t)m * D + d], box[d]);
        s = (d == 0) ? dd[d] * dd[d] : s + dd[d] * dd[d];
      }
      const double r = sqrt(s);
      const int tm = types ? types[m] : 0;
      const double sg = sigma[tn * T + tm], ep = eps[tn * T + tm];
      const double xr = r / sg;
      if (xr > r_cut) continue; /* f_cut, u_cut, c_cut are 0 there */
      const double sf = ep * pot_f(&p, xr);
      if (F)
        for (int d = 0; d < D; ++d) fn[d] = fn[d] + sf * (dd[d] / r);
      if (e_pp) e = e + ep * (pot_u(&p, xr) - p.u_rc);
      if (w_pp) w = w + sf * r;
      if (lap) l = l + (ep * pot_c(&p, xr) + (D - 1) * (-sf) / r);
    }
    if (F)
      for (int d = 0; d < D; ++d) F[n * D + d] = fn[d];
    if (e_pp) e_pp[n] = e;
    if (w_pp) w_pp[n] = w;
    if (lap) lap[n] = l;
  }
}

/* integrator.py:17-40 (integ = 1, G-JF) / :5-14 (integ = 2, leap-frog), then
 * positions.py:90-98.  xi[N][D] standard normals (G-JF only). */
void dpo_step(int64_t N, int D, double* x, double* v, double* beta, const double* F, const double* m, double dt,
              double T, double tau, const double* xi, int integ, int32_t* img, const double* box) {
#pragma omp parallel for schedule(static)
  for (int64_t n = 0; n < N; ++n) {
    const double mm = m[n];
    for (int d = 0; d < D; ++d) {
      const int64_t k = n * D + d;
      double xn;
      if (integ == 1) {
        const double alpha = mm / tau;
        const double bvar = 2 * alpha * T * dt;
        const double nb = sqrt(bvar) * xi[k];
        const double num = 1.0 - alpha * dt / 2.0 / mm;
        const double den = 1.0 + alpha * dt / 2.0 / mm;
        const double a = num / den;
        const double b = 1 / den;
        const double v1 = a * v[k] + b * dt * F[k] / mm + b * (beta[k] + nb) / 2 / mm;
        xn = x[k] + v1 * dt;
        v[k] = v1;
        beta[k] = nb;
      } else {
        const double v1 = v[k] + dt * F[k] / mm;
        xn = x[k] + dt * v1;
        v[k] = v1;
      }
      const double sh = floor(xn / box[d]);
      img[k] += (int32_t)sh;
      x[k] = xn - sh * box[d];
    }
  }
}

/* ------------------------------------------------------------------------
 * Philox4x32-10 (Salmon et al., SC'11) and Box-Muller, the same stream as the
 * device (counter = particle, dim/2, step lo, step hi; key = seed).
 * ---------------------------------------------------------------------- */
void dpo_philox4x32_10(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4]) {
  uint32_t c0 = ctr[0], c1 = ctr[1], c2 = ctr[2], c3 = ctr[3];
  uint32_t k0 = key[0], k1 = key[1];
  for (int r = 0; r < 10; ++r) {
    const uint64_t p0 = (uint64_t)0xD2511F53u * c0;
    const uint64_t p1 = (uint64_t)0xCD9E8D57u * c2;
    const uint32_t hi0 = (uint32_t)(p0 >> 32), lo0 = (uint32_t)p0;
    const uint32_t hi1 = (uint32_t)(p1 >> 32), lo1 = (uint32_t)p1;
    const uint32_t n0 = hi1 ^ c1 ^ k0, n1 = lo1, n2 = hi0 ^ c3 ^ k1, n3 = lo0;
    c0 = n0; c1 = n1; c2 = n2; c3 = n3;
    k0 += 0x9E3779B9u;
    k1 += 0xBB67AE85u;
  }
  out[0] = c0; out[1] = c1; out[2] = c2; out[3] = c3;
}

void dpo_normals(int64_t N, int D, uint64_t seed, uint64_t step, const int32_t* pid, double* xi) {
#pragma omp parallel for schedule(static)
  for (int64_t n = 0; n < N; ++n) {
    for (int d = 0; d < D; ++d) {
      uint32_t ctr[4] = {(uint32_t)(pid ? pid[n] : n), (uint32_t)(d >> 1), (uint32_t)step, (uint32_t)(step >> 32)};
      uint32_t key[2] = {(uint32_t)seed, (uint32_t)(seed >> 32)};
      uint32_t o[4];
      dpo_philox4x32_10(ctr, key, o);
      const uint64_t a = ((uint64_t)o[0] << 32) | o[1];
      const uint64_t c = ((uint64_t)o[2] << 32) | o[3];
      const double u1 = ((double)(a >> 11) + 0.5) * 1.1102230246251565404e-16;
      const double u2 = ((double)(c >> 11) + 0.5) * 1.1102230246251565404e-16;
      const double rad = sqrt(-2.0 * log(u1));
      const double th = 6.283185307179586476925 * u2;
      xi[n * D + d] = (d & 1) ? rad * sin(th) : rad * cos(th);
    }
  }
}
