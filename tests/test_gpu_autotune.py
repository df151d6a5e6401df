"""The GPU autotuner keeps dompap's tools/autotune.py interface and the
reference's check (tests/test_autotune.py:5-29: 0.1 < skin < 1.9); the tuned
kernel shape gives the same forces (FP64 mode) as the planner's."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def test_autotune_reference_case():
    from dompap_amd import Simulation
    from dompap_amd.tools import autotune
    sim = Simulation(seed=3)
    fcc = ((0.0, 0.0, 0.0), (0.5, 0.5, 0.0), (0.5, 0.0, 0.5), (0.0, 0.5, 0.5))
    sim.set_positions(unit_cell_coordinates=fcc, cells=(4, 4, 4), lattice_constants=(1.0, 1.0, 1.0))
    sim.set_density(density=0.8)
    sim.set_masses(masses=1.0)
    sim.set_random_velocities(temperature=1.0)
    sim.set_pair_potential(pair_potential_str="(1-r)**2", r_cut=1.0)
    sim.set_pair_potential_parameters(sigma=1.0, epsilon=1.0)
    sim.set_neighbor_list(skin=0.5, max_number_of_neighbors=128)
    sim.set_integrator(time_step=0.01, target_temperature=1.0, temperature_damping_time=0.1)
    sim.set_neighbor_list(skin=1.0, max_number_of_neighbors=512)
    F0 = sim.copy().get_forces()
    sim = autotune(sim, steps=10)
    assert 0.1 < sim.neighbor_list_skin < 1.9, f"{sim.neighbor_list_skin=}"
    tpp, tile, fused = sim.__dict__["_tuning"]
    assert tpp in (1, 2, 4, 8) and fused in (0, 1)
    F1 = sim.get_forces()
    assert np.abs(F1 - F0).max() <= 1e-12 * max(1.0, np.abs(F0).max())
    assert sim.get_kernel_plan()["threads_per_particle"] == tpp
