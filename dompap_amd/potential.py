"""Pair potentials: host callables with the reference's API plus the device
code and tables the GPU kernels need.

* ``make_pair_potential(str, r_cut)`` returns the three host callables of
  dompap's potential.py:43-66 (u shifted by u(r_cut); u, f = -u', u'' all 0
  where r > r_cut), evaluated with NumPy.  They serve host-side uses such as
  ``sim.pair_potential(r)`` (examples/lennard_jones.py:30); forces, energies
  and virials of a configuration are computed on the GPU.
* ``hardcoded_pair_potentials`` mirrors potential.py:31-40.
* ``lower_parameters`` turns the reference's sigma/epsilon callables
  (Simulation.py:469-525) into per-type tables for the device.
"""
from __future__ import annotations

import math
from dataclasses import dataclass

import numpy as np
import sympy as sp

from . import codegen


def _harmonic_pot(r):
    r = np.asarray(r, dtype=np.float64)
    return np.where(r < 1, 0.5 * (1 - r) ** 2, 0.0)


def _harmonic_force(r):
    r = np.asarray(r, dtype=np.float64)
    return np.where(r < 1, 1 - r, 0.0)


def _harmonic_curvature(r):
    # potential.py:26-28 returns 0.0 (not the true curvature 1.0); kept as is
    return np.zeros_like(np.asarray(r, dtype=np.float64))


hardcoded_pair_potentials = {
    # name: (pair_potential_str, energy, force, curvature, r_cut)  -- potential.py:31-40
    "Harmonic repulsive": ("0.5*(1-r)**2", _harmonic_pot, _harmonic_force, _harmonic_curvature, 1.0),
}


def make_pair_potential(pair_potential_str="(((1-r)**2)**(1/2))**(7/2)", r_cut=1.0):
    """potential.py:43-66: (u_cut, f_cut, u''_cut) host callables."""
    r = sp.symbols("r")
    u_e = sp.sympify(pair_potential_str)
    f_e = sp.diff(-u_e, r)
    c_e = sp.diff(u_e, r, 2)
    u_fn = sp.lambdify(r, u_e, "numpy")
    f_fn = sp.lambdify(r, f_e, "numpy")
    c_fn = sp.lambdify(r, c_e, "numpy")
    rc = float(r_cut)
    with np.errstate(all="ignore"):
        u_rc = np.float64(u_fn(np.float64(rc)))

    def _ev(fn, x):
        with np.errstate(all="ignore"):
            v = fn(x)
        return np.broadcast_to(np.asarray(v, dtype=np.float64), np.shape(x))

    def pair_potential_cut(x):
        x = np.asarray(x, dtype=np.float64)
        return np.where(x > rc, 0, _ev(u_fn, x) - u_rc)

    def pair_force_cut(x):
        x = np.asarray(x, dtype=np.float64)
        return np.where(x > rc, 0, _ev(f_fn, x))

    def pair_curvature_cut(x):
        x = np.asarray(x, dtype=np.float64)
        return np.where(x > rc, 0, _ev(c_fn, x))

    return pair_potential_cut, pair_force_cut, pair_curvature_cut


def shift_value(pair_potential_str: str, r_cut: float) -> float:
    """u(r_cut) as the reference subtracts it (potential.py:57-59); 0 for the
    hard-coded potential (no shift, potential.py:10-15)."""
    if pair_potential_str in hardcoded_pair_potentials:
        return 0.0
    r = sp.symbols("r")
    fn = sp.lambdify(r, sp.sympify(pair_potential_str), "numpy")
    with np.errstate(all="ignore"):
        v = float(np.float64(fn(np.float64(r_cut))))
    return v


@dataclass
class DevicePotential:
    code: codegen.PotentialCode
    r_cut: float
    u_rc: float


_CODE_CACHE: dict = {}


def device_potential(pair_potential_str: str, r_cut: float) -> DevicePotential:
    key = pair_potential_str
    if key not in _CODE_CACHE:
        _CODE_CACHE[key] = codegen.generate(pair_potential_str)
    if pair_potential_str in hardcoded_pair_potentials:
        r_cut = hardcoded_pair_potentials[pair_potential_str][4]
    return DevicePotential(_CODE_CACHE[key], float(r_cut), shift_value(pair_potential_str, float(r_cut)))


# ---------------------------------------------------------------------------
# sigma / epsilon lowering
# ---------------------------------------------------------------------------
@dataclass
class PairTables:
    types: np.ndarray      # int32 [N] dense type index per particle
    sigma: np.ndarray      # [T, T]
    epsilon: np.ndarray    # [T, T]
    sigma_max: float       # max over all (n, m) incl. n == m (Simulation.py:347-351)


def _dense_types(particle_types) -> tuple[np.ndarray, np.ndarray]:
    t = np.asarray(particle_types).reshape(-1)
    uniq, dense = np.unique(t, return_inverse=True)
    return uniq, dense.astype(np.int32)


def lower_parameters(sigma_func, epsilon_func, particle_types, sigma_const=None, epsilon_const=None,
                     rng_seed: int = 12345, n_checks: int = 4096, max_types: int = 8) -> PairTables:
    """Per-type sigma/epsilon tables for the device.

    Constant parameters give 1x1 tables.  Callables of (n, m) are evaluated on
    one representative pair per type pair; a random sample of (n, m), n != m,
    is then checked against the tables and a ValueError is raised if the
    callable depends on more than the two particle types (the GPU kernels read
    sigma/epsilon per type pair).  The diagonal sigma(n, n) only enters the
    neighbour-list cutoff (Simulation.py:347-351) and is scanned separately.
    """
    uniq, dense = _dense_types(particle_types)
    n = dense.size
    nt = len(uniq)
    if sigma_const is not None and epsilon_const is not None:
        s = float(sigma_const)
        e = float(epsilon_const)
        return PairTables(np.zeros(n, dtype=np.int32), np.full((1, 1), s), np.full((1, 1), e), s)
    if nt > max_types:
        raise ValueError(f"at most {max_types} particle types are supported on the GPU, got {nt}")
    first = [np.nonzero(dense == t)[0] for t in range(nt)]
    sig = np.zeros((nt, nt))
    eps = np.zeros((nt, nt))
    for a in range(nt):
        for b in range(nt):
            ia = first[a]
            ib = first[b]
            if a == b:
                if ia.size < 2:
                    nn, mm = int(ia[0]), int(ia[0])
                    sv = sigma_func(nn, mm) if sigma_const is None else sigma_const
                    ev = epsilon_func(nn, mm) if epsilon_const is None else epsilon_const
                    sig[a, b], eps[a, b] = float(sv), float(ev)
                    continue
                nn, mm = int(ia[0]), int(ia[1])
            else:
                nn, mm = int(ia[0]), int(ib[0])
            sig[a, b] = float(sigma_func(nn, mm)) if sigma_const is None else float(sigma_const)
            eps[a, b] = float(epsilon_func(nn, mm)) if epsilon_const is None else float(epsilon_const)
    if n > 1:
        rng = np.random.default_rng(rng_seed)
        k = min(n_checks, n * (n - 1))
        ns = rng.integers(0, n, size=k)
        ms = rng.integers(0, n - 1, size=k)
        ms = ms + (ms >= ns)
        for a, b in zip(ns.tolist(), ms.tolist()):
            ta, tb = dense[a], dense[b]
            if sigma_const is None and float(sigma_func(a, b)) != sig[ta, tb]:
                raise ValueError("sigma depends on more than the particle types of (n, m); "
                                 "assign particle types so that sigma(n, m) is a function of them")
            if epsilon_const is None and float(epsilon_func(a, b)) != eps[ta, tb]:
                raise ValueError("epsilon depends on more than the particle types of (n, m); "
                                 "assign particle types so that epsilon(n, m) is a function of them")
    smax = float(np.max(sig)) if sigma_const is None else float(sigma_const)
    if sigma_const is None:
        if n <= 200_000:
            diag = max(float(sigma_func(i, i)) for i in range(n))
        else:
            idx = np.unique(np.concatenate([f[:64] for f in first] +
                                           [np.random.default_rng(rng_seed).integers(0, n, 4096)]))
            diag = max(float(sigma_func(int(i), int(i))) for i in idx)
        smax = max(smax, diag, 0.0)
    if not (np.all(sig > 0) and math.isfinite(smax)):
        raise ValueError("sigma must be positive and finite")
    return PairTables(dense, sig, eps, smax)
