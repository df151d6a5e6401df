"""The GPU autotuner keeps dompap's tools/autotune.py interface and the
reference's check (tests/test_autotune.py:5-29: 0.1 < skin < 1.9); the tuned
kernel shape, cell width and row capacity give the same forces (FP64 mode)
as the planner's, per system shape."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def test_autotune_reference_case():
    from dompap_amd import Simulation
    from dompap_amd.tools import autotune
    from dompap_amd.tools.autotune import CELL_SCALES
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
    assert sim.__dict__["_cell_scale"] in CELL_SCALES
    # row capacity: lowered from 512 to 1.5x the longest row, in multiples of 8
    K = int(sim.max_number_of_neighbors)
    F1 = sim.get_forces()
    longest = int(np.max(np.sum(np.asarray(sim.neighbor_list) >= 0, axis=1)))
    assert K % 8 == 0 and 64 <= K < 512 and K >= longest
    assert np.abs(F1 - F0).max() <= 1e-12 * max(1.0, np.abs(F0).max())
    assert sim.get_kernel_plan()["threads_per_particle"] == tpp


def test_autotune_lj_liquid_picks_shape_cells_and_capacity():
    """A 3-D LJ liquid (N = 32,000): the tuned plan (lanes per particle,
    cell width, capacity) runs and gives the planner's forces."""
    from dompap_amd import Simulation
    from dompap_amd.tools import autotune
    sim = Simulation(seed=5)
    fcc = ((0.0, 0.0, 0.0), (0.5, 0.5, 0.0), (0.5, 0.0, 0.5), (0.0, 0.5, 0.5))
    sim.set_positions(unit_cell_coordinates=fcc, cells=(20, 20, 20))
    sim.set_density(density=0.8442)
    sim.set_masses(masses=1.0)
    sim.set_random_velocities(temperature=0.72)
    sim.set_pair_potential(pair_potential_str="4*(r**-12-r**-6)", r_cut=2.5)
    sim.set_neighbor_list(skin=0.5, max_number_of_neighbors=512)
    sim.set_integrator(time_step=0.005, target_temperature=0.72, temperature_damping_time=10.0)
    sim.run(50)
    F0 = sim.copy().get_forces()
    sim = autotune(sim, steps=20, smallest_skin=0.3, step_skin=0.2)
    F1 = sim.get_forces()
    frms = float(np.sqrt(np.mean(np.sum(F0 ** 2, axis=1))))
    assert np.abs(F1 - F0).max() / frms < 1e-12
    plan = sim.get_kernel_plan()
    assert plan["threads_per_particle"] == sim.__dict__["_tuning"][0]
    assert 64 <= int(sim.max_number_of_neighbors) <= 512


def test_cell_width_tuning_keeps_lists_and_forces():
    """tune_cell_scale (the bench's --cell-scale 0) times each cell width on
    the engine's own system and keeps the fastest: wider cells change the
    tiles and the emit's candidate tables, not the neighbour sets or the
    forces (FP64 mode, a 3-D LJ liquid of 32,000 particles)."""
    from dompap_amd.engine import Engine
    from dompap_amd.potential import device_potential
    from dompap_amd.tools.autotune import CELL_SCALES, tune_cell_scale
    from oracle import ref_numpy as R
    fcc = ((0.0, 0.0, 0.0), (0.5, 0.5, 0.0), (0.5, 0.0, 0.5), (0.0, 0.5, 0.5))
    a = (4 / 0.8442) ** (1 / 3)
    x = R.generate_positions(fcc, (20, 20, 20), (a, a, a))
    box = np.full(3, 20 * a)
    rng = np.random.default_rng(2)
    x = (x + rng.normal(0, 0.05, size=x.shape)) % box
    eng = Engine(3, x.shape[0], 1, "fp64")
    eng.set_box(box)
    dp = device_potential("4*(r**-12-r**-6)", 2.5)
    eng.load_potential(dp.code, dp.r_cut, dp.u_rc, [[1.0]], [[1.0]])
    eng.set_neighbor_params(3.0, 0.5, 256)
    eng.set_integrator(0.005, 0.72, 10.0)
    eng.set_seed(4, 0)
    eng.set_field("positions", x)
    eng.set_field("velocities", rng.normal(0, 0.85, size=x.shape))
    eng.build_neighbor_list()
    scale, times = tune_cell_scale(eng, steps=40)
    assert scale in CELL_SCALES and set(times) == set(CELL_SCALES)
    xs = eng.get_field("positions")
    eng.build_neighbor_list()
    rows = [np.sort(r[r >= 0]) for r in eng.neighbor_list()]
    F = eng.compute(forces=True)["forces"]
    eng.set_cell_scale(1.0)
    eng.set_field("positions", xs)
    eng.build_neighbor_list()
    rows1 = [np.sort(r[r >= 0]) for r in eng.neighbor_list()]
    F1 = eng.compute(forces=True)["forces"]
    assert all(np.array_equal(p, q) for p, q in zip(rows, rows1))
    frms = float(np.sqrt(np.mean(np.sum(F1 ** 2, axis=1))))
    assert np.abs(F - F1).max() / frms < 1e-12
    eng.close()
