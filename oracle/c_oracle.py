"""ctypes wrapper of the C oracle (oracle/libdp_oracle.so) -- TEST INFRASTRUCTURE.

Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg use it.
Potential ids: "lj" = 4*(r**-12-r**-6), "harm2" = (1-r)**2,
"harmonic" = the reference's hard-coded "Harmonic repulsive", ("ipl", n) = r**-n.
"""
from __future__ import annotations

import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(HERE, "libdp_oracle.so")

_P = {"lj": 0, "harm2": 1, "harmonic": 2, "ipl": 3}
_lib = None


def load():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB):
            from . import build_oracle
            build_oracle.build(quiet=True)
        _lib = ctypes.CDLL(LIB)
        vp = ctypes.c_void_p
        i64, ci, cd = ctypes.c_int64, ctypes.c_int, ctypes.c_double
        _lib.dpo_nl_double_loop.argtypes = [i64, ci, vp, vp, cd, ci, vp, vp]
        _lib.dpo_nl_double_loop.restype = ci
        _lib.dpo_nl_cells.argtypes = [i64, ci, vp, vp, cd, ci, vp, vp]
        _lib.dpo_nl_cells.restype = ci
        _lib.dpo_max_displacement2.argtypes = [i64, ci, vp, vp, vp]
        _lib.dpo_max_displacement2.restype = cd
        _lib.dpo_pair.argtypes = [i64, ci, vp, vp, vp, ci, ci, cd, cd, ci, vp, vp, vp, vp, vp, vp, vp]
        _lib.dpo_pair.restype = None
        _lib.dpo_step.argtypes = [i64, ci, vp, vp, vp, vp, vp, cd, cd, cd, vp, ci, vp, vp]
        _lib.dpo_step.restype = None
        _lib.dpo_normals.argtypes = [i64, ci, ctypes.c_uint64, ctypes.c_uint64, vp, vp]
        _lib.dpo_normals.restype = None
        _lib.dpo_philox4x32_10.argtypes = [vp, vp, vp]
        _lib.dpo_philox4x32_10.restype = None
        _lib.dpo_num_threads.restype = ci
        _lib.dpo_set_threads.argtypes = [ci]
        _lib.dpo_shift.argtypes = [ci, cd, cd]
        _lib.dpo_shift.restype = cd
    return _lib


def _p(a):
    return None if a is None else a.ctypes.data_as(ctypes.c_void_p)


def _pot(pot):
    if isinstance(pot, tuple):
        return _P[pot[0]], float(pot[1])
    return _P[pot], 0.0


def threads() -> int:
    return load().dpo_num_threads()


def set_threads(n: int) -> None:
    load().dpo_set_threads(int(n))


def neighbor_list(x, box, cutoff, K, method="auto"):
    """neighbor_list.py:7-39 rows (ascending m, -1 padded) and row lengths."""
    lib = load()
    x = np.ascontiguousarray(x, dtype=np.float64)
    box = np.ascontiguousarray(box, dtype=np.float64)
    N, D = x.shape
    nl = np.empty((N, K), dtype=np.int32)
    cnt = np.empty(N, dtype=np.int32)
    if method == "auto":
        method = "double loop" if N <= 20000 else "cells"
    fn = lib.dpo_nl_double_loop if method == "double loop" else lib.dpo_nl_cells
    fn(N, D, _p(x), _p(box), float(cutoff), int(K), _p(nl), _p(cnt))
    return nl, cnt


def max_displacement2(x, xref, box) -> float:
    x = np.ascontiguousarray(x, dtype=np.float64)
    xref = np.ascontiguousarray(xref, dtype=np.float64)
    box = np.ascontiguousarray(box, dtype=np.float64)
    return load().dpo_max_displacement2(x.shape[0], x.shape[1], _p(x), _p(xref), _p(box))


def pair(x, box, nl, pot, r_cut, sigma=1.0, eps=1.0, types=None, forces=True, energy=False, virial=False,
         laplacian=False):
    """Forces [N,D], energy (potential.py:100 halving), virial (/2/D over the
    full list) and per-particle Laplacian."""
    lib = load()
    x = np.ascontiguousarray(x, dtype=np.float64)
    box = np.ascontiguousarray(box, dtype=np.float64)
    nl = np.ascontiguousarray(nl, dtype=np.int32)
    N, D = x.shape
    sig = np.atleast_2d(np.asarray(sigma, dtype=np.float64))
    ep = np.atleast_2d(np.asarray(eps, dtype=np.float64))
    T = max(sig.shape[0], ep.shape[0])  # a scalar table applies to every type pair
    sig = np.ascontiguousarray(np.broadcast_to(sig, (T, T)))
    ep = np.ascontiguousarray(np.broadcast_to(ep, (T, T)))
    ty = None if types is None else np.ascontiguousarray(types, dtype=np.int32).reshape(-1)
    F = np.empty((N, D)) if forces else None
    E = np.empty(N) if energy else None
    W = np.empty(N) if virial else None
    L = np.empty(N) if laplacian else None
    pid, pn = _pot(pot)
    lib.dpo_pair(N, D, _p(x), _p(box), _p(nl), nl.shape[1], pid, pn, float(r_cut), T, _p(ty), _p(sig), _p(ep),
                 _p(F), _p(E), _p(W), _p(L))
    out = {"forces": F}
    if energy:
        out["energy"] = float(np.sum(E)) / 2
    if virial:
        out["virial"] = float(np.sum(W)) / 2 / D
    if laplacian:
        out["laplacian"] = L
    return out


def step(x, v, beta, F, masses, dt, T, tau, xi, box, img, integrator="gjf"):
    """integrator.py:17-40 / :5-14 + wrap, in place on copies; returns them."""
    lib = load()
    x = np.array(x, dtype=np.float64, order="C")
    v = np.array(v, dtype=np.float64, order="C")
    beta = np.array(beta, dtype=np.float64, order="C")
    img = np.array(img, dtype=np.int32, order="C")
    F = np.ascontiguousarray(F, dtype=np.float64)
    m = np.ascontiguousarray(np.asarray(masses, dtype=np.float64).reshape(-1))
    box = np.ascontiguousarray(box, dtype=np.float64)
    xi = None if xi is None else np.ascontiguousarray(xi, dtype=np.float64)
    N, D = x.shape
    lib.dpo_step(N, D, _p(x), _p(v), _p(beta), _p(F), _p(m), float(dt), float(T), float(tau), _p(xi),
                 1 if integrator == "gjf" else 2, _p(img), _p(box))
    return x, v, beta, img


def normals(N, D, seed, step_index, pid=None):
    """The device's Philox4x32-10 / Box-Muller noise for one step."""
    xi = np.empty((N, D))
    p = None if pid is None else np.ascontiguousarray(pid, dtype=np.int32)
    load().dpo_normals(N, D, int(seed) & (2**64 - 1), int(step_index), _p(p), _p(xi))
    return xi


def philox4x32_10(ctr, key):
    c = np.ascontiguousarray(ctr, dtype=np.uint32)
    k = np.ascontiguousarray(key, dtype=np.uint32)
    o = np.empty(4, dtype=np.uint32)
    load().dpo_philox4x32_10(_p(c), _p(k), _p(o))
    return o
