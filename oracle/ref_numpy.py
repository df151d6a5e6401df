"""FP64 NumPy restatement of dompap's per-step hot path -- TEST INFRASTRUCTURE.

This module is the oracle.  It is not the product: only ``tests/``,
``__graft_entry__.smoke()`` and the ``cpu_baseline`` leg of ``bench.py`` may
import it, and only to check results.  The product (``dompap_amd``) never
imports anything under ``oracle/``.

Every function restates one reference function, with the file:line it follows
(paths relative to the dompap repository).  Arithmetic is written in the same
operation order as the reference (no fused multiply-add, sequential sums) so
that neighbour sets and single-step results can be compared bit for bit where
the reference's own semantics allow it.

Pinning (what checks each function; see tests/test_oracle.py):

* ``neighbor_list_double_loop``: tests/test_neighbor_list.py:12-51 row KATs,
  the Simulation.py:327-333 doctest (max 26 neighbours, strict ``<``), and a
  brute-force distance-matrix comparison on random inputs.
* ``neighbor_list_cell_list`` (row order only): tests/test_neighbor_list.py:54-91.
* ``neighbor_list_is_old``: tests/test_neighbor_list.py:94-118.
* ``make_pair_potential``: tests/test_potential.py:20-39 and closed forms.
* ``total_energy`` / ``forces`` / ``virial_double_loop`` / ``laplacian``:
  tests/test_potential.py:81-140,191-224 KATs, doctests Simulation.py:551-555
  (E=167.06442443573454), :600-605 (F0=(-1,0,0)), :952-955 (W=0),
  :1211-1215 (lap0=12.0), the reference-held fixture
  examples/harmonic_repulsive_3d.{csv,toml} (W=5.022787227389184),
  closed-form lattice sums, Newton's third law and the inverse-power-law
  identity W = (n/D) U.
* ``laplacian``'s first-derivative term (D - 1)(-eps f)/r: central second
  differences of ``total_energy`` in D = 2, 3, 4.
* ``make_one_step`` / ``make_one_step_leap_frog``: closed forms of
  arXiv:1303.7011 Eqs. 16-17 with injected noise, the energy-conservation
  bound of tests/test_integrator.py:6-104, and at finite tau three
  properties the mathematics fixes: the T = 0 terminal velocity F/alpha
  (the b dt F/m term), the free-particle stationary <m v^2> =
  T/(1 + dt/2tau) (the b (beta + beta')/2m noise term) and <kappa x^2> = T
  in a harmonic well.
* ``wrap_into_box`` / ``minimage``: tests/test_positions.py:12-25,78-93.
* ``rdf_counts``: closed-form shell counts of simple-cubic and fcc lattices
  (per particle 1 at r = 0, then 6, 12, 8, 6 and 12, 6, 24, 12), the
  ideal-gas limit g(r) ~ 1 on tests/test_positions.py:66-76's case, and
  equality with the unblocked positions.get_radial_distribution_function.
"""
from __future__ import annotations

import math

import numpy as np
import sympy as sp


# --------------------------------------------------------------------------
# positions.py
# --------------------------------------------------------------------------
def displacement(a, b, box):
    """positions.py:6-17 -- d = a - b, then one +-L shift per dimension
    (``> L/2`` tested first)."""
    d = np.array(a, dtype=np.float64) - np.array(b, dtype=np.float64)
    for k in range(d.shape[-1]):
        half = box[k] / 2
        dk = d[..., k]
        dk = np.where(dk > half, dk - box[k], np.where(dk < -half, dk + box[k], dk))
        d[..., k] = dk
    return d


def distance(a, b, box):
    """positions.py:20-30 -- sqrt of the sequential sum of squares."""
    d = displacement(a, b, box)
    s = d[..., 0] * d[..., 0]
    for k in range(1, d.shape[-1]):
        s = s + d[..., k] * d[..., k]
    return np.sqrt(s)


def generate_positions(unit_cell_coordinates, cells, lattice_constants):
    """positions.py:33-52 -- lattice in C order (last cell index fastest),
    unit-cell sites innermost, then scaled by the lattice constants."""
    ucc = np.asarray(unit_cell_coordinates, dtype=np.float64)
    cells = np.asarray(cells, dtype=np.int64)
    lc = np.asarray(lattice_constants, dtype=np.float64)
    ncell = int(np.prod(cells))
    idx = np.array(np.unravel_index(np.arange(ncell), tuple(cells)), dtype=np.float64).T
    # reference: positions[...] = ucc + cell_coordinates (ucc first)
    pos = (ucc[None, :, :] + idx[:, None, :]).reshape(-1, ucc.shape[1])
    return pos * lc


def rdf_counts(positions, box, r_bins, block=512):
    """positions.py:58-87 without the normalisation -- the np.histogram
    counts of the minimum-image distances of every ordered pair, self pairs
    included: d = x_n - x_m; d -= L * np.round(d / L); r = np.linalg.norm(d)
    (NumPy 2.2 sums the squares in dimension order).  Rows are processed in
    blocks; the per-pair arithmetic is the reference's, and histogram counts
    of blocks add up exactly."""
    x = np.asarray(positions, dtype=np.float64)
    r_bins = np.asarray(r_bins, dtype=np.float64)
    out = np.zeros(r_bins.size - 1, dtype=np.int64)
    for i0 in range(0, x.shape[0], block):
        diff = x[i0:i0 + block, np.newaxis, :] - x[np.newaxis, :, :]
        for i in range(diff.shape[-1]):
            box_length = np.linalg.norm(box[i])
            diff[:, :, i] -= box_length * np.round(diff[:, :, i] / box_length)
        out += np.histogram(np.linalg.norm(diff, axis=2), bins=r_bins)[0]
    return out


def wrap_into_box(positions, image_positions, box):
    """positions.py:90-98 -- s = floor(x/L); img += int32(s); x -= s*L (in place)."""
    shift = np.floor(positions / box[np.newaxis, :])
    image_positions += shift.astype(np.int32)
    positions -= shift * box[np.newaxis, :]


# --------------------------------------------------------------------------
# neighbor_list.py
# --------------------------------------------------------------------------
def _sum_sq_sequential(d):
    s = d[..., 0] * d[..., 0]
    for k in range(1, d.shape[-1]):
        s = s + d[..., k] * d[..., k]
    return s


def neighbor_list_double_loop(positions, box, cutoff, max_neighbors):
    """neighbor_list.py:7-39 -- full list, ascending m != n, strict
    ``sqrt(sum d^2) < cutoff``; min-image tests ``< -L/2`` first (:27-31);
    -1 padded.  Unlike the reference, overflow raises ValueError instead of
    writing out of bounds (neighbor_list.py:36-38)."""
    x = np.asarray(positions, dtype=np.float64)
    n_part, dim = x.shape
    nl = np.full((n_part, int(max_neighbors)), -1, dtype=np.int32)
    for n in range(n_part):
        d = x[n][None, :] - x
        for k in range(dim):
            half = box[k] / 2
            dk = d[:, k]
            d[:, k] = np.where(dk < -half, dk + box[k], np.where(dk > half, dk - box[k], dk))
        r = np.sqrt(_sum_sq_sequential(d))
        hit = r < cutoff
        hit[n] = False
        ms = np.nonzero(hit)[0]
        if ms.size > max_neighbors:
            raise ValueError(f"particle {n} has {ms.size} neighbours > max_number_of_neighbors={max_neighbors}")
        nl[n, : ms.size] = ms
    return nl


def neighbor_list_cell_list(positions, box, cutoff, max_neighbors, x_fastest=None):
    """Row order of neighbor_list.py:42-103 (any D, C-order stencil via
    ``unravel_index``, last dimension fastest) and of :106-180 (3-D,
    x fastest, :150-157).  Cells are ``ceil(L/cutoff)`` wide-``cutoff`` bins
    filled in ascending particle order, exactly as the reference; like the
    reference this misses pairs when L/cutoff is not an integer and repeats
    cells when fewer than three fit, so it is only used for row-order KATs."""
    x = np.asarray(positions, dtype=np.float64)
    n_part, dim = x.shape
    if x_fastest is None:
        x_fastest = dim == 3
    ncell = np.ceil(box / cutoff).astype(np.int64)
    cells = {}
    for n in range(n_part):
        c = tuple(np.floor(x[n] / cutoff).astype(np.int64))
        cells.setdefault(c, []).append(n)
    nl = np.full((n_part, int(max_neighbors)), -1, dtype=np.int32)
    for n in range(n_part):
        c = np.floor(x[n] / cutoff).astype(np.int64)
        k = 0
        for i in range(3 ** dim):
            if x_fastest:
                shift = np.array([(i // 3 ** d) % 3 - 1 for d in range(dim)])
            else:
                shift = np.array(np.unravel_index(i, (3,) * dim)) - 1
            nc = c + shift
            for d in range(dim):
                if nc[d] < 0:
                    nc[d] += ncell[d]
                elif nc[d] >= ncell[d]:
                    nc[d] -= ncell[d]
            for m in cells.get(tuple(nc), []):
                if distance(x[n], x[m], box) < cutoff and m != n:
                    nl[n, k] = m
                    k += 1
    return nl


def neighbor_list_is_old(positions, old_positions, box, skin):
    """neighbor_list.py:183-201 -- max_n |minimage(x - x_ref)|^2 > (skin/2)^2."""
    d = displacement(positions, old_positions, box)
    s = _sum_sq_sequential(d)
    mx = 0.0
    if s.size:
        mx = max(0.0, float(np.max(s)))
    return bool(mx > (0.5 * skin) ** 2)


def number_of_neighbors(nl):
    """neighbor_list.py:204-207."""
    return np.sum(nl != -1, axis=1)


def neighbor_sets(nl):
    """Row -> frozenset of neighbour ids (order-free comparison helper)."""
    return [frozenset(int(m) for m in row if m != -1) for row in nl]


# --------------------------------------------------------------------------
# potential.py
# --------------------------------------------------------------------------
HARDCODED = {
    # potential.py:10-40: name -> (expression string, r_cut); force 1-r and a
    # curvature that returns 0.0 (sic), strict r < 1 support.
    "Harmonic repulsive": ("0.5*(1-r)**2", 1.0),
}


def make_pair_potential(pair_potential_str="(1-r)**2", r_cut=1.0):
    """potential.py:43-66 -- u = sympify(str); f = d(-u)/dr; u'' = d^2u/dr^2;
    lambdify(numpy); u shifted by u(r_cut); all three cut to 0 where r > r_cut
    (``>``: r == r_cut is kept)."""
    if pair_potential_str in HARDCODED:
        def u(r):
            r = np.asarray(r, dtype=np.float64)
            return np.where(r < 1, 0.5 * (1 - r) ** 2, 0.0)

        def f(r):
            r = np.asarray(r, dtype=np.float64)
            return np.where(r < 1, 1 - r, 0.0)

        def c(r):
            return np.zeros_like(np.asarray(r, dtype=np.float64))

        return u, f, c
    r = sp.symbols("r")
    u_expr = sp.sympify(pair_potential_str)
    f_expr = sp.diff(-u_expr, r)
    c_expr = sp.diff(u_expr, r, 2)
    u_fn = sp.lambdify(r, u_expr, "numpy")
    f_fn = sp.lambdify(r, f_expr, "numpy")
    c_fn = sp.lambdify(r, c_expr, "numpy")
    rc = float(r_cut)
    with np.errstate(all="ignore"):
        u_rc = np.float64(u_fn(np.float64(rc))) if math.isfinite(rc) else np.float64(u_fn(np.float64(rc)))

    def _arr(fn, x):
        with np.errstate(all="ignore"):
            v = fn(np.asarray(x, dtype=np.float64))
        return np.broadcast_to(np.asarray(v, dtype=np.float64), np.shape(x))

    def u_cut(x):
        x = np.asarray(x, dtype=np.float64)
        return np.where(x > rc, 0.0, _arr(u_fn, x) - u_rc)

    def f_cut(x):
        x = np.asarray(x, dtype=np.float64)
        return np.where(x > rc, 0.0, _arr(f_fn, x))

    def c_cut(x):
        x = np.asarray(x, dtype=np.float64)
        return np.where(x > rc, 0.0, _arr(c_fn, x))

    return u_cut, f_cut, c_cut


def _pair_table(func, n, m):
    """Evaluate a sigma/epsilon callable on index arrays (scalar calls)."""
    return np.array([func(int(a), int(b)) for a, b in zip(np.atleast_1d(n), np.atleast_1d(m))],
                    dtype=np.float64)


def _row(positions, box, n, ms):
    d = positions[n][None, :] - positions[ms]
    for k in range(positions.shape[1]):
        half = box[k] / 2
        dk = d[:, k]
        d[:, k] = np.where(dk > half, dk - box[k], np.where(dk < -half, dk + box[k], dk))
    return d


def total_energy(positions, box, pair_potential, nl, sigma_func, epsilon_func):
    """potential.py:69-100 -- U = (sum over rows, in row order, of
    eps * u(r / sigma)) / 2, accumulated sequentially."""
    x = np.asarray(positions, dtype=np.float64)
    terms = []
    for n in range(x.shape[0]):
        row = nl[n]
        ms = row[: np.argmax(row == -1)] if np.any(row == -1) else row
        if ms.size == 0:
            continue
        d = _row(x, box, n, ms)
        r = np.sqrt(_sum_sq_sequential(d))
        sig = _pair_table(sigma_func, np.full(ms.size, n), ms)
        eps = _pair_table(epsilon_func, np.full(ms.size, n), ms)
        terms.append(eps * pair_potential(r / sig))
    if not terms:
        return 0.0
    t = np.concatenate(terms)
    return float(np.cumsum(t)[-1] / 2)


def forces(positions, box, pair_force, nl, sigma_func, epsilon_func):
    """potential.py:163-188 -- F_n = sum_{m in NL(n)} eps * f(r/sigma) * d / r,
    d = minimage(x_n - x_m), r = (sum d^2)**0.5; no 1/sigma factor (sic)."""
    x = np.asarray(positions, dtype=np.float64)
    out = np.zeros_like(x)
    for n in range(x.shape[0]):
        row = nl[n]
        ms = row[: np.argmax(row == -1)] if np.any(row == -1) else row
        if ms.size == 0:
            continue
        d = _row(x, box, n, ms)
        r = _sum_sq_sequential(d) ** 0.5
        sig = _pair_table(sigma_func, np.full(ms.size, n), ms)
        eps = _pair_table(epsilon_func, np.full(ms.size, n), ms)
        sf = eps * pair_force(r / sig)
        contrib = sf[:, None] * (d / r[:, None])
        acc = np.zeros(x.shape[1])
        for c in contrib:  # sequential, row order
            acc = acc + c
        out[n] = acc
    return out


def virial_double_loop(positions, box, pair_force, sigma_func, epsilon_func):
    """potential.py:253-279 -- W = (1/D) sum_{n<m} eps * f(r/sigma) * r over
    all pairs."""
    x = np.asarray(positions, dtype=np.float64)
    n_part, dim = x.shape
    w = 0.0
    for n in range(n_part - 1):
        ms = np.arange(n + 1, n_part)
        d = _row(x, box, n, ms)
        r = _sum_sq_sequential(d) ** 0.5
        sig = _pair_table(sigma_func, np.full(ms.size, n), ms)
        eps = _pair_table(epsilon_func, np.full(ms.size, n), ms)
        w += float(np.sum(eps * pair_force(r / sig) * r))
    return float(w / dim)


def virial_neighbor_list(positions, box, pair_force, nl, sigma_func, epsilon_func):
    """Same quantity as ``virial_double_loop`` summed over a full neighbour
    list and halved (every pair with f != 0 is in a list built with cutoff
    >= sigma_max * r_cut)."""
    x = np.asarray(positions, dtype=np.float64)
    w = 0.0
    for n in range(x.shape[0]):
        row = nl[n]
        ms = row[: np.argmax(row == -1)] if np.any(row == -1) else row
        if ms.size == 0:
            continue
        d = _row(x, box, n, ms)
        r = _sum_sq_sequential(d) ** 0.5
        sig = _pair_table(sigma_func, np.full(ms.size, n), ms)
        eps = _pair_table(epsilon_func, np.full(ms.size, n), ms)
        w += float(np.sum(eps * pair_force(r / sig) * r))
    return float(w / 2 / x.shape[1])


def laplacian(positions, box, pair_force, pair_curvature, nl, sigma_func, epsilon_func):
    """potential.py:327-364 -- per particle sum of eps*u''(r/sigma)
    + (D-1) * (-eps*f(r/sigma)) / r (no sigma factors, as the reference)."""
    x = np.asarray(positions, dtype=np.float64)
    dim = x.shape[1]
    out = np.zeros(x.shape[0])
    for n in range(x.shape[0]):
        row = nl[n]
        ms = row[: np.argmax(row == -1)] if np.any(row == -1) else row
        if ms.size == 0:
            continue
        d = _row(x, box, n, ms)
        r = _sum_sq_sequential(d) ** 0.5
        sig = _pair_table(sigma_func, np.full(ms.size, n), ms)
        eps = _pair_table(epsilon_func, np.full(ms.size, n), ms)
        dv = -(eps * pair_force(r / sig))
        d2v = eps * pair_curvature(r / sig)
        terms = d2v + (dim - 1) * dv / r
        out[n] = float(np.cumsum(terms)[-1])
    return out


# --------------------------------------------------------------------------
# integrator.py
# --------------------------------------------------------------------------
def make_one_step(positions, velocities, betas, forces_, masses, time_step,
                  temperature_target, temperature_damping_time, xi):
    """integrator.py:17-40 (G-JF, arXiv:1303.7011 Eqs. 16-17) with the
    standard normals ``xi`` supplied by the caller instead of Numba's
    internal generator (integrator.py:30)."""
    alphas = masses / temperature_damping_time
    beta_variance = 2 * alphas * temperature_target * time_step
    new_betas = np.sqrt(beta_variance) * xi
    numerator = 1.0 - alphas * time_step / 2.0 / masses
    denominator = 1.0 + alphas * time_step / 2.0 / masses
    a = numerator / denominator
    b = 1 / denominator
    velocities = a * velocities + b * time_step * forces_ / masses + b * (betas + new_betas) / 2 / masses
    positions = positions + velocities * time_step
    return positions, velocities, new_betas


def make_one_step_leap_frog(positions, velocities, forces_, masses, time_step):
    """integrator.py:5-14."""
    velocities = velocities + time_step * forces_ / masses
    positions = positions + time_step * velocities
    return positions, velocities


# --------------------------------------------------------------------------
# Simulation.py orchestration
# --------------------------------------------------------------------------
def nl_cutoff(sigma_max, r_cut, skin):
    """Simulation.py:346-356 -- cutoff = max sigma * r_cut + skin."""
    return sigma_max * r_cut + skin


class OracleSim:
    """Minimal restatement of Simulation.step / update_neighbor_list
    (Simulation.py:307-401, 681-724) for trajectory comparisons, with noise
    supplied by the caller."""

    def __init__(self, positions, box, velocities, masses, pot_str, r_cut,
                 skin, max_neighbors=512, sigma=1.0, epsilon=1.0,
                 dt=0.01, T=1.0, tau=0.1, betas=None):
        self.x = np.array(positions, dtype=np.float64)
        self.box = np.array(box, dtype=np.float64)
        self.v = np.array(velocities, dtype=np.float64)
        self.m = np.array(masses, dtype=np.float64).reshape(-1, 1)
        self.beta = np.zeros_like(self.x) if betas is None else np.array(betas, dtype=np.float64)
        self.img = np.zeros(self.x.shape, dtype=np.int32)
        self.u, self.f, self.c = make_pair_potential(pot_str, r_cut)
        self.r_cut = HARDCODED[pot_str][1] if pot_str in HARDCODED else float(r_cut)
        self.skin = float(skin)
        self.K = int(max_neighbors)
        self.sigma = sigma if callable(sigma) else (lambda n, m, s=float(sigma): s)
        self.epsilon = epsilon if callable(epsilon) else (lambda n, m, e=float(epsilon): e)
        self.dt, self.T, self.tau = float(dt), float(T), float(tau)
        self.updates = 0
        self.steps = 0
        self.set_neighbor_list()

    def set_neighbor_list(self):
        self.updates += 1
        n = self.x.shape[0]
        smax = 0.0
        for a in range(n):
            for b in range(n):
                smax = max(smax, self.sigma(a, b))
        self.cutoff = nl_cutoff(smax, self.r_cut, self.skin)
        self.nl = neighbor_list_double_loop(self.x, self.box, self.cutoff, self.K)
        self.x_ref = self.x.copy()

    def update_neighbor_list(self):
        if neighbor_list_is_old(self.x, self.x_ref, self.box, self.skin):
            self.set_neighbor_list()

    def get_forces(self):
        self.update_neighbor_list()
        return forces(self.x, self.box, self.f, self.nl, self.sigma, self.epsilon)

    def get_potential_energy(self):
        self.update_neighbor_list()
        return total_energy(self.x, self.box, self.u, self.nl, self.sigma, self.epsilon)

    def get_virial(self):
        return virial_double_loop(self.x, self.box, self.f, self.sigma, self.epsilon)

    def step(self, xi):
        self.update_neighbor_list()
        F = self.get_forces()
        self.x, self.v, self.beta = make_one_step(self.x, self.v, self.beta, F, self.m,
                                                  self.dt, self.T, self.tau, xi)
        wrap_into_box(self.x, self.img, self.box)
        self.steps += 1

    def step_leap_frog(self):
        self.update_neighbor_list()
        F = self.get_forces()
        self.x, self.v = make_one_step_leap_frog(self.x, self.v, F, self.m, self.dt)
        wrap_into_box(self.x, self.img, self.box)
        self.steps += 1
