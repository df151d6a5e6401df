"""Betas (integrator.py:30-31, beta' = sqrt(2 alpha T dt) xi) across step
kinds and state changes.

Split steps (force kernel, then integrate kernel) and fused steps (one
kernel) run the same arithmetic: over the same sequence of step blocks,
single steps, leap-frog steps, and integrator and mass changes
they must agree bit for bit.  After G-JF steps with device noise the betas
are the oracle's Philox normals of the last G-JF step times its noise
amplitude; leap-frog steps keep them; host-set betas replace them."""
import math

import numpy as np
import pytest

from oracle import c_oracle as C
from oracle import ref_numpy as R

pytestmark = pytest.mark.gpu
FCC = [(0, 0, 0), (.5, .5, 0), (.5, 0, .5), (0, .5, .5)]


def _lj(precision, fused):
    from dompap_amd import Simulation
    s = Simulation(seed=11, precision=precision)
    s.set_positions(unit_cell_coordinates=FCC, cells=(5, 5, 5))
    s.set_density(0.9)
    s.set_masses(1.0)
    rng = np.random.default_rng(5)
    s.positions = np.asarray(s.positions) + rng.normal(0, 0.05, size=s.positions.shape)
    s.set_pair_potential("4*(r**-12-r**-6)", r_cut=2.5)
    s.set_neighbor_list(skin=0.5, max_number_of_neighbors=256)
    s.set_integrator(time_step=0.005, target_temperature=1.0, temperature_damping_time=1.0)
    s.set_kernel_tuning(2, 0, fused)
    s.velocities = rng.normal(0, 1.0, size=s.positions.shape)
    s.betas = np.zeros(s.positions.shape)
    return s


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_unstored_betas_match_stored_betas(precision):
    runs = []
    for fused in (0, 1):
        s = _lj(precision, fused)
        out = []
        s.run(37)                      # step blocks, rebuilds
        out.append(np.asarray(s.betas).copy())
        s.step()
        s.step()                       # single steps
        s.step_leap_frog()             # keeps the last G-JF step's betas
        s.step_leap_frog()
        out.append(np.asarray(s.betas).copy())
        s.run(20)
        s.set_integrator(target_temperature=2.0)  # a new noise amplitude
        s.run(15)
        s.set_masses(2.0)              # amplitudes per mass
        s.run(15)
        for name in ("positions", "velocities", "betas", "image_positions"):
            out.append(np.asarray(getattr(s, name)).copy())
        runs.append(out)
    for k, (a, b) in enumerate(zip(*runs)):
        assert np.array_equal(a, b), k


def test_unstored_betas_are_the_last_steps_noise():
    from dompap_amd.engine import Engine
    from dompap_amd.potential import device_potential
    n = 1000
    eng = Engine(3, n, 1, "mixed")
    eng.set_box([20.0, 20.0, 20.0])
    dp = device_potential("(1-r)**2", 1.0)
    eng.load_potential(dp.code, 1.0, dp.u_rc, [[1.0]], [[1.0]])
    eng.set_neighbor_params(1.5, 0.5, 64)
    dt, T, tau = 0.01, 2.0, 0.5
    eng.set_integrator(dt, T, tau)
    eng.set_seed(12345, 77)
    x = R.generate_positions([[0, 0, 0]], (10, 10, 10), (2.0, 2.0, 2.0))
    eng.set_field("positions", x)
    eng.build_neighbor_list()
    amp = math.sqrt(2 * (1.0 / tau) * T * dt)
    # 40 G-JF steps (step blocks): the betas of the last one, index 77 + 39
    eng.step(40, "gjf")
    assert np.abs(eng.get_field("betas") - amp * C.normals(n, 3, 12345, 116)).max() < 1e-13
    # leap-frog steps 117 .. 119 keep them
    eng.step(3, "lf")
    assert np.abs(eng.get_field("betas") - amp * C.normals(n, 3, 12345, 116)).max() < 1e-13
    # G-JF steps 120 .. 124, then a new temperature: the betas keep the old amplitude
    eng.step(5, "gjf")
    eng.set_integrator(dt, 2 * T, tau)
    assert np.abs(eng.get_field("betas") - amp * C.normals(n, 3, 12345, 124)).max() < 1e-13
    eng.step(1, "gjf")
    amp2 = math.sqrt(2 * (1.0 / tau) * 2 * T * dt)
    assert np.abs(eng.get_field("betas") - amp2 * C.normals(n, 3, 12345, 125)).max() < 1e-13
    # betas set by the host replace unstored ones
    b = np.full((n, 3), 0.25)
    eng.step(2, "gjf")
    eng.set_field("betas", b)
    assert np.array_equal(eng.get_field("betas"), b)
    eng.close()
