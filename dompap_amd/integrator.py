"""Integrator free functions with the reference's API (dompap/integrator.py),
run by the device integrator kernel (libdompap_hip dp_integrate) with the
caller's forces.  As in the reference, positions are not wrapped here."""
from __future__ import annotations

import numpy as np

from .engine import Engine
from .potential import device_potential


def _engine(positions, velocities, betas, masses, time_step, temperature_target, temperature_damping_time):
    x = np.ascontiguousarray(np.asarray(positions, dtype=np.float64))
    n, dim = x.shape
    eng = Engine(dim, n, 1, "fp64")
    eng.set_box(np.ones(dim))
    dp = device_potential("(1-r)**2", 1.0)
    eng.load_potential(dp.code, dp.r_cut, dp.u_rc, [[1.0]], [[1.0]])
    eng.set_neighbor_params(0.0, 1.0, 1)
    eng.set_integrator(float(time_step), float(temperature_target), float(temperature_damping_time))
    eng.set_field("positions", x)
    eng.set_field("velocities", np.asarray(velocities, dtype=np.float64))
    if betas is not None:
        eng.set_field("betas", np.asarray(betas, dtype=np.float64))
    eng.set_field("masses", np.asarray(masses, dtype=np.float64).reshape(-1))
    return eng


def make_one_step_leap_frog(positions, velocities, forces, masses, time_step):
    """integrator.py:5-14."""
    eng = _engine(positions, velocities, None, masses, time_step, 1.0, np.inf)
    try:
        eng.integrate(np.asarray(forces, dtype=np.float64), "leapfrog", wrap=False)
        return eng.get_field("positions"), eng.get_field("velocities")
    finally:
        eng.close()


def make_one_step(positions, velocities, betas, forces, masses, time_step, temperature_target,
                  temperature_damping_time, noise=None):
    """integrator.py:17-40 (G-JF).  ``noise`` supplies the standard normals
    of integrator.py:30; by default they are drawn from numpy.random."""
    if noise is None:
        noise = np.random.normal(0, 1, np.shape(velocities))
    eng = _engine(positions, velocities, betas, masses, time_step, temperature_target, temperature_damping_time)
    try:
        eng.integrate(np.asarray(forces, dtype=np.float64), "gjf", xi=np.asarray(noise, dtype=np.float64),
                      wrap=False)
        return eng.get_field("positions"), eng.get_field("velocities"), eng.get_field("betas")
    finally:
        eng.close()
