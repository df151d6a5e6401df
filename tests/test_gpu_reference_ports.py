"""The reference's own integration tests, run on the GPU through dompap_amd:

* tests/test_lennard_jones.py:69-152 -- the LJ crystal at coexistence
  (J. Chem. Phys. 139, 104102 (2013), Table I): <T>, <T_conf>, <K>/N within
  10 %, <P> = 2.185 within 10 %, E/N = -4.953 and U/N within 0.5;
* tests/test_inverse_power_law.py:58-127 -- r^-12: the W-versus-U slope
  4 +- 0.9, intercept 0 +- 0.9, correlation 1 +- 0.9;
* tests/test_to_and_from_disk.py:9-62 -- the same non-default state written
  and read back, here checked field by field (bit for bit);

plus the reference rows no other test runs on the GPU: the hard-coded
"Harmonic repulsive" potential (potential.py:10-40, with its zero
curvature), the alternate force / energy / list method strings
(Simulation.py:126-133), and get_configurational_temperature
(Simulation.py:838-856) against the oracle.

Seeds are fixed (np.random.seed for the reference's global-state velocity
draw, Simulation(seed=...) for the device noise); tolerances are the
reference's."""
import os

import numpy as np
import pytest

from oracle import ref_numpy as R

pytestmark = pytest.mark.gpu
FCC = np.array([(0.0, 0.0, 0.0), (0.5, 0.5, 0.0), (0.5, 0.0, 0.5), (0.0, 0.5, 0.5)], dtype=np.float64)
ONE = lambda n, m: 1.0  # noqa: E731


def _run(sim, steps, stride, target_temperature, tau, conf=True):
    """run_simulation of the reference tests: set_integrator(dt 0.004, T, tau),
    then `steps` steps with the observables every `stride` steps."""
    sim.set_integrator(time_step=0.004, target_temperature=target_temperature, temperature_damping_time=tau)
    rows = []
    for step in range(steps):
        if step % stride == 0:
            N = sim.get_number_of_particles()
            row = [sim.get_time(), sim.get_potential_energy() / N, sim.get_temperature(),
                   sim.get_kinetic_energy() / N, sim.get_virial() / N, sim.get_pressure()]
            if conf:
                row.append(sim.get_configurational_temperature())
            rows.append(row)
        sim.step()
    return np.asarray(rows)


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_lennard_jones_crystal(precision):
    # tests/test_lennard_jones.py:69-152
    from dompap_amd import Simulation
    np.random.seed(2013)
    sim = Simulation(seed=104102, precision=precision)
    sim.set_positions(unit_cell_coordinates=FCC, cells=(6, 6, 6))
    sim.set_density(density=1.0 / 1.0277)
    sim.set_masses(masses=1.0)
    sim.set_particle_types(types=0)
    sim.set_random_velocities(temperature=0.8 * 2)
    sim.set_pair_potential(pair_potential_str="4*(r**-12-r**-6)", r_cut=2.5)
    sim.set_pair_potential_parameters(sigma=1.0, epsilon=1.0)
    sim.set_neighbor_list(skin=1.2, max_number_of_neighbors=512)
    sim.set_integrator(time_step=0.004, target_temperature=0.8, temperature_damping_time=2.0)
    _run(sim, 200, 1, 0.8, 0.1)   # equilibrate
    df = _run(sim, 200, 1, 0.8, 0.1)  # production
    U, T, K, W, P, Tc = (df[:, k] for k in range(1, 7))
    assert 0.8 * 0.9 < T.mean() < 0.8 * 1.1, T.mean()
    assert 0.8 * 0.9 < Tc.mean() < 0.8 * 1.1, Tc.mean()
    assert 1.2 * 0.9 < K.mean() < 1.2 * 1.1, K.mean()
    assert 2.185 * 0.9 < P.mean() < 2.185 * 1.1, P.mean()
    u_exp = -4.953 - 0.8 * 3 / 2
    assert u_exp - 0.5 < U.mean() < u_exp + 0.5, U.mean()
    assert -4.953 - 0.5 < U.mean() + K.mean() < -4.953 + 0.5
    # the pressure the reference recomputes from T, density and W
    p_re = T * sim.get_density() + W * sim.get_number_of_particles() / sim.get_volume()
    assert np.allclose(p_re, P, rtol=1e-12, atol=1e-12)


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_inverse_power_law(precision):
    # tests/test_inverse_power_law.py:58-127
    from dompap_amd import Simulation
    np.random.seed(12)
    sim = Simulation(seed=4, precision=precision)
    sim.set_positions(unit_cell_coordinates=FCC, cells=(6, 6, 6))
    sim.set_density(density=0.7)
    sim.set_masses(masses=1.0)
    sim.set_particle_types(types=0)
    sim.set_random_velocities(temperature=1.0)
    sim.set_pair_potential(pair_potential_str="r**-12", r_cut=2.0)
    sim.set_pair_potential_parameters(sigma=1.0, epsilon=1.0)
    sim.set_neighbor_list(skin=1.2, max_number_of_neighbors=512)
    sim.set_integrator(time_step=0.004, target_temperature=1.0, temperature_damping_time=2.0)
    _run(sim, 40, 4, 0.8, 2.0, conf=False)          # equilibrate (run_simulation's default T = 0.8)
    df = _run(sim, 40, 4, 0.8, 2.0, conf=False)     # production
    x, y = df[:, 1], df[:, 4]
    slope = np.cov(x, y)[0, 1] / np.var(x)
    intersection = np.mean(y) - slope * np.mean(x)
    corr = np.corrcoef(x, y)[0, 1]
    assert 4 - 0.9 < slope < 4 + 0.9, slope
    assert -0.9 < intersection < 0.9, intersection
    assert 1 - 0.9 < corr < 1 + 0.9, corr


def test_to_and_from_disk_round_trip(tmp_path):
    # tests/test_to_and_from_disk.py:9-62 (box 4.2 x 4 x 4.8, list cutoff
    # 1.1 * 1.2 + 0.54321: fewer than three cells along each side)
    from dompap_amd import Simulation
    np.random.seed(3)
    sim = Simulation(seed=5)
    sim.set_positions(unit_cell_coordinates=((0, 0, 0), (0.5, 0.5, 0.5)), cells=(1, 2, 2),
                      lattice_constants=(4.2, 2.0, 2.4))
    sim.set_masses(masses=1.234)
    sim.set_random_velocities(temperature=1.0)
    sim.set_pair_potential(pair_potential_str="(1.2-r)**4", r_cut=1.2)
    sim.set_pair_potential_parameters(sigma=1.1, epsilon=0.24)
    sim.set_neighbor_list(skin=0.54321, max_number_of_neighbors=192)
    sim.set_integrator(time_step=0.1234, target_temperature=1.123, temperature_damping_time=0.123)
    sim.particle_types[0] = 1
    sim.particle_types[1] = 2
    sim.image_positions[0] = 1, 2, 3
    sim.image_positions[1] = 4, 5, 6
    sim.betas[0] = 1.2, 3.4, 5.6
    sim.betas[1] = 7.8, 9.1, 2.3
    sim.run(7)  # device-owned state (positions, velocities, betas, images) goes through the files
    pcsv, ptoml = os.path.join(tmp_path, "simulation.csv"), os.path.join(tmp_path, "simulation.toml")
    sim.to_disk(particle_data=pcsv, meta_data=ptoml)
    new = Simulation(seed=5)
    meta = new.from_disk(particle_data=pcsv, meta_data=ptoml)
    for name in ("positions", "velocities", "betas", "image_positions", "masses", "particle_types"):
        a, b = np.asarray(getattr(sim, name)), np.asarray(getattr(new, name))
        assert a.dtype == b.dtype and a.shape == b.shape and np.array_equal(a, b), name
    assert np.array_equal(np.asarray(sim.box_vectors), np.asarray(new.box_vectors))
    for key in ("pair_potential_str", "pair_potential_r_cut", "neighbor_list_skin", "max_number_of_neighbors",
                "neighbor_list_method_str", "energy_method_str", "force_method_str", "time_step",
                "temperature_target", "temperature_damping_time", "number_of_steps",
                "number_of_neighbor_list_updates"):
        assert meta[key] == getattr(new, key) or float(meta[key]) == float(getattr(new, key)), key
    assert meta["number_of_steps"] == 7
    assert meta["potential_energy"] == sim.get_potential_energy()
    assert meta["virial"] == sim.get_virial()
    # the reloaded state evaluates to the same forces (sigma / epsilon are
    # not part of the files: set them as the original had them)
    new.set_pair_potential_parameters(sigma=1.1, epsilon=0.24)
    assert np.array_equal(new.get_forces(), sim.get_forces())
    # and steps on identically (same seed, same step counter)
    sim.run(3)
    new.run(3)
    assert np.array_equal(np.asarray(new.positions), np.asarray(sim.positions))


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_hardcoded_harmonic_repulsive(precision):
    # potential.py:10-40: 0.5 (1 - r)^2, force 1 - r, curvature 0.0 (sic), r_cut 1
    from dompap_amd import Simulation
    sim = Simulation(seed=1, precision=precision)
    sim.set_positions(unit_cell_coordinates=([0.0, 0.0, 0.0],), cells=(8, 8, 8), lattice_constants=(0.95,) * 3)
    rng = np.random.default_rng(2)
    x = np.asarray(sim.positions) + rng.normal(0, 0.06, (512, 3))
    sim.positions = x
    sim.set_pair_potential("Harmonic repulsive")
    sim.set_neighbor_list(skin=0.3, max_number_of_neighbors=64)
    assert sim.pair_potential_r_cut == 1.0
    box = np.asarray(sim.box_vectors)
    u, f, c = R.make_pair_potential("Harmonic repulsive", 1.0)
    nl = R.neighbor_list_double_loop(x, box, 1.3, 64)
    assert np.array_equal(R.neighbor_sets(sim.neighbor_list), R.neighbor_sets(nl))
    Fo = R.forces(x, box, f, nl, ONE, ONE)
    frms = float(np.sqrt(np.mean(np.sum(Fo ** 2, axis=1))))
    tol = 1e-10 if precision == "fp64" else 1e-5
    assert np.abs(sim.get_forces() - Fo).max() / frms < tol
    Eo = R.total_energy(x, box, u, nl, ONE, ONE)
    assert abs(sim.get_potential_energy() - Eo) < (1e-10 if precision == "fp64" else 1e-6) * abs(Eo)
    Wo = R.virial_double_loop(x, box, f, ONE, ONE)
    assert abs(sim.get_virial() - Wo) < (1e-10 if precision == "fp64" else 1e-6) * abs(Wo)
    lap = sim.get_laplacian()
    Lo = R.laplacian(x, box, f, c, nl, ONE, ONE)
    assert np.abs(lap - Lo).max() < (1e-10 if precision == "fp64" else 1e-5) * np.abs(Lo).max()
    # zero curvature: only the first-derivative term (D - 1)(-f)/r, which is
    # negative for every overlapping pair
    assert np.all(lap <= 0.0) and np.any(lap < 0.0)
    # the hard-coded form differs from the SymPy string "(1-r)**2" by the 1/2
    sim2 = Simulation(seed=1, precision=precision)
    sim2.set_positions(unit_cell_coordinates=([0.0, 0.0, 0.0],), cells=(8, 8, 8), lattice_constants=(0.95,) * 3)
    sim2.positions = x
    sim2.set_pair_potential("(1-r)**2", r_cut=1.0)
    sim2.set_neighbor_list(skin=0.3, max_number_of_neighbors=64)
    assert np.allclose(sim2.get_forces(), 2 * sim.get_forces(), rtol=1e-12, atol=1e-14)


def test_alternate_method_strings_give_the_same_results():
    # Simulation.py:126-133 method strings: every force / energy method and
    # both list methods give the default results (same physics, one set of
    # kernels); the list rows come out ascending for "double loop"
    # (neighbor_list.py:21-35), and "cell list" holds the same sets
    from dompap_amd import Simulation
    sim = Simulation(seed=3)
    sim.set_positions(unit_cell_coordinates=FCC, cells=(5, 5, 5))
    sim.set_density(0.9)
    rng = np.random.default_rng(4)
    x = np.asarray(sim.positions) + rng.normal(0, 0.05, (500, 3))
    sim.positions = x
    sim.set_pair_potential("4*(r**-12-r**-6)", r_cut=2.5)
    sim.set_neighbor_list(skin=0.4, max_number_of_neighbors=200)
    F0, E0, W0 = sim.get_forces(), sim.get_potential_energy(), sim.get_virial()
    for fm in sorted(Simulation._KNOWN_FORCE_METHODS):
        for em in sorted(Simulation._KNOWN_ENERGY_METHODS):
            sim.set_pair_potential("4*(r**-12-r**-6)", r_cut=2.5, force_method=fm, energy_method=em)
            assert sim.force_method_str == fm and sim.energy_method_str == em
            assert np.array_equal(sim.get_forces(), F0), fm
            assert sim.get_potential_energy() == E0, em
            assert sim.get_virial() == W0
    with pytest.raises(ValueError):
        sim.set_pair_potential("4*(r**-12-r**-6)", r_cut=2.5, force_method="no such method")
    with pytest.raises(ValueError):
        sim.set_pair_potential("4*(r**-12-r**-6)", r_cut=2.5, energy_method="no such method")
    box = np.asarray(sim.box_vectors)
    nl_ref = R.neighbor_list_double_loop(x, box, 2.9, 200)
    sim.set_neighbor_list(method_str="double loop")
    assert np.array_equal(sim.neighbor_list, nl_ref)
    sim.set_neighbor_list(method_str="cell list")
    assert R.neighbor_sets(sim.neighbor_list) == R.neighbor_sets(nl_ref)
    assert np.array_equal(sim.get_forces(), F0)
    with pytest.raises(ValueError):
        sim.set_neighbor_list(method_str="no such method")


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_configurational_temperature(precision):
    # Simulation.py:838-856: T_conf = <|F|^2> / <lap U> over the particles
    from dompap_amd import Simulation
    sim = Simulation(seed=3, precision=precision)
    sim.set_positions(unit_cell_coordinates=FCC, cells=(5, 5, 5))
    sim.set_density(0.95)
    rng = np.random.default_rng(6)
    x = np.asarray(sim.positions) + rng.normal(0, 0.06, (500, 3))
    sim.positions = x
    sim.set_pair_potential("4*(r**-12-r**-6)", r_cut=2.5)
    sim.set_neighbor_list(skin=0.5, max_number_of_neighbors=200)
    box = np.asarray(sim.box_vectors)
    u, f, c = R.make_pair_potential("4*(r**-12-r**-6)", 2.5)
    nl = R.neighbor_list_double_loop(x, box, 3.0, 200)
    Fo = R.forces(x, box, f, nl, ONE, ONE)
    Lo = R.laplacian(x, box, f, c, nl, ONE, ONE)
    t_ref = float(np.mean(np.sum(Fo ** 2, axis=1)) / np.mean(Lo))
    t = sim.get_configurational_temperature()
    assert abs(t - t_ref) <= (1e-10 if precision == "fp64" else 1e-5) * abs(t_ref), (t, t_ref)
    assert t > 0
