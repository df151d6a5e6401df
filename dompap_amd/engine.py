"""Engine: one libdompap_hip context (one or more replicas on one GPU).

Arrays cross the boundary in the reference's layout (row n = particle n).
Inputs may be NumPy arrays (host) or torch device tensors (their data_ptr()
is handed to the library; it copies device-to-device).  Outputs are NumPy
arrays unless an ``out`` tensor is given.
"""
from __future__ import annotations

import ctypes
import math
from typing import Optional, Sequence

import numpy as np

from . import _lib
from .codegen import PotentialCode


def _ptr(a):
    """Raw pointer of a contiguous NumPy array or torch tensor."""
    if hasattr(a, "data_ptr"):
        if not a.is_contiguous():
            raise ValueError("tensor must be contiguous")
        return ctypes.c_void_p(a.data_ptr())
    return a.ctypes.data_as(ctypes.c_void_p)


def _as_f64(a, shape):
    if hasattr(a, "data_ptr"):
        import torch  # noqa: F401  (only for device tensors)
        if a.dtype != __import__("torch").float64 or tuple(a.shape) != tuple(shape):
            raise ValueError(f"expected float64 tensor of shape {shape}")
        return a.contiguous()
    arr = np.ascontiguousarray(np.asarray(a, dtype=np.float64))
    if arr.shape != tuple(shape):
        arr = arr.reshape(shape)
    return arr


def _as_i32(a, shape):
    if hasattr(a, "data_ptr"):
        if tuple(a.shape) != tuple(shape):
            raise ValueError(f"expected tensor of shape {shape}")
        return a.contiguous()
    arr = np.ascontiguousarray(np.asarray(a, dtype=np.int32))
    if arr.shape != tuple(shape):
        arr = arr.reshape(shape)
    return arr


class Engine:
    """Device state and kernels for ``n_total`` particles in ``n_replicas``
    independent systems of dimension ``dim``."""

    def __init__(self, dim: int, n_total: int, n_replicas: int = 1, precision: str = "fp64", device: int = 0):
        self.lib = _lib.load()
        if precision not in ("fp64", "mixed"):
            raise ValueError("precision must be 'fp64' or 'mixed'")
        self.dim = int(dim)
        self.n = int(n_total)
        self.n_replicas = int(n_replicas)
        self.precision = precision
        ctx = ctypes.c_void_p()
        rc = self.lib.dp_create(ctypes.byref(ctx), int(device), self.dim, self.n, self.n_replicas,
                                _lib.DP_MIXED if precision == "mixed" else _lib.DP_FP64)
        _lib.check(rc, None)
        self.ctx = ctx
        self.max_neighbors = 0

    # -- lifetime ---------------------------------------------------------
    def close(self):
        if getattr(self, "ctx", None):
            self.lib.dp_destroy(self.ctx)
            self.ctx = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _chk(self, rc):
        _lib.check(rc, self.ctx)

    # -- parameters -------------------------------------------------------
    def _per_rep(self, v):
        a = np.ascontiguousarray(np.broadcast_to(np.asarray(v, dtype=np.float64), (self.n_replicas,)))
        return a

    def set_replicas(self, offsets: Sequence[int], box) -> None:
        off = np.ascontiguousarray(np.asarray(offsets, dtype=np.int64))
        bx = np.ascontiguousarray(np.asarray(box, dtype=np.float64).reshape(self.n_replicas, self.dim))
        self._chk(self.lib.dp_set_replicas(self.ctx, _ptr(off), _ptr(bx)))

    def set_box(self, box) -> None:
        """Single-replica convenience."""
        self.set_replicas([0, self.n], np.asarray(box, dtype=np.float64).reshape(1, self.dim))

    def set_integrator(self, dt, temperature, damping_time) -> None:
        a, b, c = self._per_rep(dt), self._per_rep(temperature), self._per_rep(damping_time)
        self._chk(self.lib.dp_set_integrator(self.ctx, _ptr(a), _ptr(b), _ptr(c)))

    def set_neighbor_params(self, cutoff, skin, max_neighbors: int, order: int = _lib.DP_ORDER_X_FASTEST) -> None:
        a, b = self._per_rep(cutoff), self._per_rep(skin)
        self._chk(self.lib.dp_set_neighbor_params(self.ctx, _ptr(a), _ptr(b), int(max_neighbors), int(order)))
        self.max_neighbors = int(max_neighbors)

    def load_potential(self, code: PotentialCode, r_cut: float, u_rc: float, sigma, epsilon,
                       pair_precision: str = "fp64") -> None:
        """pair_precision (mixed mode only): "fp64" evaluates in-range pair
        terms in FP64 (default), "fp32" in FP32; the candidate screening and
        the displacements are fixed-point / FP32 in both cases."""
        sig = np.ascontiguousarray(np.atleast_2d(np.asarray(sigma, dtype=np.float64)))
        eps = np.ascontiguousarray(np.atleast_2d(np.asarray(epsilon, dtype=np.float64)))
        nt = sig.shape[0]
        if sig.shape != (nt, nt) or eps.shape != (nt, nt):
            raise ValueError("sigma/epsilon tables must be square and of equal size")
        acc = {"fp64": 0, "fp32": 1}[pair_precision]
        rc = self.lib.dp_load_potential(self.ctx, code.source.encode(), int(code.even), float(r_cut),
                                        float(u_rc), nt, _ptr(sig), _ptr(eps), acc)
        self._chk(rc)

    def load_potentials(self, codes, r_cuts, u_rcs, sigmas, epsilons, rep_pot, pair_precision: str = "fp64") -> None:
        """Several potentials (PotentialCode, r_cut, u(r_cut), sigma and
        epsilon tables of equal size), replica r using potential rep_pot[r]."""
        npot = len(codes)
        sig = np.ascontiguousarray(np.stack([np.atleast_2d(np.asarray(t, dtype=np.float64)) for t in sigmas]))
        eps = np.ascontiguousarray(np.stack([np.atleast_2d(np.asarray(t, dtype=np.float64)) for t in epsilons]))
        nt = sig.shape[1]
        if sig.shape != (npot, nt, nt) or eps.shape != (npot, nt, nt):
            raise ValueError("one square sigma / epsilon table of equal size per potential")
        srcs = (ctypes.c_char_p * npot)(*[c.source.encode() for c in codes])
        even = np.ascontiguousarray(np.array([int(c.even) for c in codes], dtype=np.int32))
        rc = np.ascontiguousarray(np.asarray(r_cuts, dtype=np.float64).reshape(npot))
        urc = np.ascontiguousarray(np.asarray(u_rcs, dtype=np.float64).reshape(npot))
        rp = np.ascontiguousarray(np.asarray(rep_pot, dtype=np.int32).reshape(self.n_replicas))
        acc = {"fp64": 0, "fp32": 1}[pair_precision]
        self._chk(self.lib.dp_load_potentials(self.ctx, npot, srcs, _ptr(even), _ptr(rc), _ptr(urc), nt, _ptr(sig),
                                              _ptr(eps), acc, _ptr(rp)))

    def set_seed(self, seed: int, step_counter: int = 0) -> None:
        self._chk(self.lib.dp_set_seed(self.ctx, int(seed) & (2**64 - 1), int(step_counter)))

    def set_replica_seeds(self, seeds) -> None:
        a = np.ascontiguousarray(np.asarray(seeds, dtype=np.uint64).reshape(self.n_replicas))
        self._chk(self.lib.dp_set_replica_seeds(self.ctx, _ptr(a)))

    def set_tuning(self, threads_per_particle: int = 0, tile_particles: int = 0, fused: int = -1) -> None:
        """Kernel shape: lanes per particle, home particles per tile (0 =
        planner's choice for both) and fused (1) / split (0) step kernels
        (-1 = planner's choice)."""
        self._chk(self.lib.dp_set_tuning(self.ctx, int(threads_per_particle), int(tile_particles), int(fused)))

    def set_cell_scale(self, scale: float) -> None:
        """Cells at least ``scale`` x cutoff wide (1 <= scale <= 4)."""
        self._chk(self.lib.dp_set_cell_scale(self.ctx, float(scale)))

    PLAN_KEYS = ("threads_per_particle", "entry_bits", "offset_bits", "fixed_point_bits", "lds_records",
                 "lds_records_emit", "tiles", "tile_grid", "fused", "home_cells_per_tile", "halo_cells_per_tile",
                 "cells", "steps_per_block", "graph_events", "entry_shift", "sentinel_padding")

    def plan(self) -> dict:
        out = np.zeros(len(self.PLAN_KEYS), dtype=np.int32)
        self._chk(self.lib.dp_get_plan(self.ctx, _ptr(out), len(out)))
        return dict(zip(self.PLAN_KEYS, (int(v) for v in out)))

    def set_graph_block(self, steps: int) -> None:
        self._chk(self.lib.dp_set_graph_block(self.ctx, int(steps)))

    def set_block_rebuild(self, mode: int = -1) -> None:
        """Step blocks rebuild lists once at their start (0) or before every
        step, per replica (1); -1 chooses 1 for several replicas."""
        self._chk(self.lib.dp_set_block_rebuild(self.ctx, int(mode)))

    def prepare_step_graphs(self, integrator: str = "gjf", timed: bool = False) -> None:
        """Capture the step-block graphs of both buffer parities now (not at
        the first step() that replays them)."""
        integ = _lib.DP_INTEGRATOR_GJF if integrator == "gjf" else _lib.DP_INTEGRATOR_LEAPFROG
        self._chk(self.lib.dp_prepare_step_graphs(self.ctx, integ, int(bool(timed))))

    # -- state ------------------------------------------------------------
    def _field(self, name):
        code, kind, wide = _lib.FIELDS[name]
        shape = (self.n, self.dim) if wide else (self.n,)
        return code, kind, shape

    def set_field(self, name: str, value) -> None:
        code, kind, shape = self._field(name)
        arr = _as_f64(value, shape) if kind == "f8" else _as_i32(value, shape)
        self._chk(self.lib.dp_set_field(self.ctx, code, _ptr(arr)))

    def get_field(self, name: str, out=None):
        code, kind, shape = self._field(name)
        if out is None:
            out = np.empty(shape, dtype=np.float64 if kind == "f8" else np.int32)
        self._chk(self.lib.dp_get_field(self.ctx, code, _ptr(out)))
        return out

    # -- neighbour list ---------------------------------------------------
    def build_neighbor_list(self) -> None:
        self._chk(self.lib.dp_build_neighbor_list(self.ctx))

    def update_neighbor_list(self) -> bool:
        rb = ctypes.c_int(0)
        self._chk(self.lib.dp_update_neighbor_list(self.ctx, ctypes.byref(rb)))
        return bool(rb.value)

    def neighbor_list(self) -> np.ndarray:
        out = np.empty((self.n, self.max_neighbors), dtype=np.int32)
        self._chk(self.lib.dp_get_neighbor_list(self.ctx, _ptr(out)))
        return out

    def neighbor_counts(self) -> np.ndarray:
        out = np.empty((self.n,), dtype=np.int32)
        self._chk(self.lib.dp_get_neighbor_counts(self.ctx, _ptr(out)))
        return out

    def stats(self):
        upd = np.zeros(self.n_replicas, dtype=np.int64)
        steps = ctypes.c_int64()
        mc = ctypes.c_int32()
        self._chk(self.lib.dp_get_stats(self.ctx, upd.ctypes.data_as(ctypes.POINTER(ctypes.c_int64)),
                                        ctypes.byref(steps), ctypes.byref(mc)))
        return upd, int(steps.value), int(mc.value)

    # -- compute ----------------------------------------------------------
    def compute(self, forces: bool = True, energy: bool = False, virial: bool = False,
                laplacian: bool = False, forces_out=None):
        F = None
        if forces:
            F = forces_out if forces_out is not None else np.empty((self.n, self.dim), dtype=np.float64)
        E = np.empty(self.n_replicas) if energy else None
        W = np.empty(self.n_replicas) if virial else None
        LAP = np.empty(self.n) if laplacian else None
        rc = self.lib.dp_compute(self.ctx,
                                 _ptr(F) if F is not None else None,
                                 _ptr(E) if E is not None else None,
                                 _ptr(W) if W is not None else None,
                                 _ptr(LAP) if LAP is not None else None)
        self._chk(rc)
        return {"forces": F, "energy": E, "virial": W, "laplacian": LAP}

    def step(self, nsteps: int = 1, integrator: str = "gjf", xi=None) -> None:
        integ = _lib.DP_INTEGRATOR_GJF if integrator == "gjf" else _lib.DP_INTEGRATOR_LEAPFROG
        xp = None
        if xi is not None:
            xi = _as_f64(xi, (self.n, self.dim))
            xp = _ptr(xi)
        self._chk(self.lib.dp_step(self.ctx, int(nsteps), integ, xp))

    def integrate(self, forces, integrator: str = "gjf", xi=None, wrap: bool = True) -> None:
        integ = _lib.DP_INTEGRATOR_GJF if integrator == "gjf" else _lib.DP_INTEGRATOR_LEAPFROG
        f = _as_f64(forces, (self.n, self.dim))
        xp = None
        if xi is not None:
            xi = _as_f64(xi, (self.n, self.dim))
            xp = _ptr(xi)
        self._chk(self.lib.dp_integrate(self.ctx, _ptr(f), integ, xp, int(bool(wrap))))

    def sync(self) -> None:
        self._chk(self.lib.dp_sync(self.ctx))

    def set_kernel_timing(self, enable: bool, every: int = 1) -> None:
        """HIP-event timing of the step kernels; with every = k > 1, of every
        step of one step block in k (the first one included)."""
        self._chk(self.lib.dp_set_kernel_timing(self.ctx, int(every) if enable else 0))

    def kernel_time(self):
        ms = ctypes.c_double()
        n = ctypes.c_int64()
        self._chk(self.lib.dp_get_kernel_time(self.ctx, ctypes.byref(ms), ctypes.byref(n)))
        return float(ms.value), int(n.value)

    def rdf_counts(self, edges) -> np.ndarray:
        """Pair-distance histogram of the (single) replica over bin edges."""
        e = np.ascontiguousarray(np.asarray(edges, dtype=np.float64))
        out = np.zeros(max(e.size - 1, 0), dtype=np.uint64)
        self._chk(self.lib.dp_rdf_counts(self.ctx, _ptr(e), int(e.size), _ptr(out)))
        return out

    def rdf_counts_noalloc(self, edges) -> np.ndarray:
        """rdf_counts through the library's allocation-free entry point (at
        most 4096 bins; the list must have been built)."""
        e = np.ascontiguousarray(np.asarray(edges, dtype=np.float64))
        out = np.zeros(max(e.size - 1, 0), dtype=np.uint64)
        self._chk(self.lib.dp_rdf_counts_noalloc(self.ctx, _ptr(e), int(e.size), _ptr(out)))
        return out

    def rdf_path(self) -> int:
        """1 if the last rdf_counts ran on the cell tiles, 0 for all pairs."""
        v = ctypes.c_int32(-1)
        self._chk(self.lib.dp_get_rdf_path(self.ctx, ctypes.byref(v)))
        return int(v.value)

    def kernel_breakdown(self):
        """(force-kernel ms, integrate-kernel ms) summed over the timed steps."""
        f = ctypes.c_double()
        g = ctypes.c_double()
        self._chk(self.lib.dp_get_kernel_breakdown(self.ctx, ctypes.byref(f), ctypes.byref(g)))
        return float(f.value), float(g.value)

    def stream(self) -> int:
        s = ctypes.c_void_p()
        self._chk(self.lib.dp_get_stream(self.ctx, ctypes.byref(s)))
        return int(s.value or 0)


def exact_is_finite(x: float) -> bool:
    return math.isfinite(float(x))
