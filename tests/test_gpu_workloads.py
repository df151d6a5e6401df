"""Full-size parity cases for the workloads other than the benchmark's:
the 3-D binary LJ mixture (N = 262,144, per-type sigma, masses 2:1), 4-D
(1-r)**2 (N = 2^21) and 2-D r**-12 (N = 2^22), all in mixed precision against the C oracle (FP64) on the same positions:
neighbour sets bit-exact, max |dF| / F_rms <= 1e-5, energy and virial
<= 1e-6 relative; then steps at scale stay finite and the rebuilt lists stay
exact."""
import numpy as np
import pytest

from oracle import c_oracle as C

pytestmark = pytest.mark.gpu


def _frms(F):
    return float(np.sqrt(np.mean(np.sum(F ** 2, axis=1))))


def _sorted_rows(nl):
    big = np.iinfo(np.int32).max
    k = np.where(nl < 0, big, nl)
    k.sort(axis=1)
    return k


def _check_against_oracle(s, pot, r_cut, cutoff, sigma=1.0, types=None):
    box = np.asarray(s.box_vectors, dtype=np.float64)
    K = int(s.max_number_of_neighbors)
    # the list is the one built at the last rebuild (positions_neighbour_list)
    nlg = np.asarray(s.neighbor_list)
    xref = np.asarray(s.positions_neighbour_list)
    nlo, cnt = C.neighbor_list(xref, box, cutoff, K, "cells")
    assert np.array_equal(_sorted_rows(nlg), _sorted_rows(nlo))
    # forces at the current positions (every pair within r_cut is listed)
    x = np.asarray(s.positions)
    nlx, _ = C.neighbor_list(x, box, cutoff, K, "cells")
    o = C.pair(x, box, nlx, pot, r_cut, sigma=sigma, types=types, energy=True, virial=True)
    F = s.get_forces()
    assert np.abs(F - o["forces"]).max() / _frms(o["forces"]) < 1e-5
    E = s.get_potential_energy()
    W = s.get_virial()
    assert abs(E - o["energy"]) / abs(o["energy"]) < 1e-6
    assert abs(W - o["virial"]) / abs(o["virial"]) < 1e-6


def test_binary_lj_mixture_262k():
    from dompap_amd import Simulation
    rng = np.random.default_rng(2)
    s = Simulation(seed=22, precision="mixed")
    a = (1 / 0.75) ** (1 / 3)
    s.set_positions(unit_cell_coordinates=([0, 0, 0],), cells=(64, 64, 64), lattice_constants=(a, a, a))
    n = s.number_of_particles()
    types = np.zeros(n, dtype=np.int32)
    types[rng.permutation(n)[: n // 2]] = 1
    s.set_types(types.tolist())
    s.masses = np.where(types == 0, 2.0, 1.0).reshape(-1, 1)
    sig = np.array([[1.0, (1 + 1 / 1.2) / 2], [(1 + 1 / 1.2) / 2, 1 / 1.2]])
    s.positions = np.asarray(s.positions) + rng.normal(0, 0.05, size=s.positions.shape)
    s.velocities = rng.normal(0, 1, size=s.positions.shape) * np.sqrt(0.7 / np.asarray(s.masses))
    s.set_pair_potential("4*(r**-12-r**-6)", r_cut=2.5)
    s.set_pair_potential_parameters(sigma=lambda i, j: sig[types[i], types[j]], epsilon=1.0)
    s.set_neighbor_list(skin=0.5, max_number_of_neighbors=192)
    s.set_integrator(time_step=0.005, target_temperature=0.7, temperature_damping_time=10.0)
    cutoff = 1.0 * 2.5 + 0.5  # sigma_max * r_cut + skin (Simulation.py:353-356)
    _check_against_oracle(s, "lj", 2.5, cutoff, sigma=sig, types=types)
    s.run(100)
    assert np.all(np.isfinite(np.asarray(s.positions)))
    _check_against_oracle(s, "lj", 2.5, cutoff, sigma=sig, types=types)


def test_4d_harmonic_repulsive_2m():
    from dompap_amd import Simulation
    rng = np.random.default_rng(4)
    s = Simulation(seed=44, precision="mixed")
    a = (1 / 1.2) ** (1 / 4)
    s.set_positions(unit_cell_coordinates=([0, 0, 0, 0],), cells=(32, 32, 32, 64), lattice_constants=(a,) * 4)
    assert s.number_of_particles() == 2 ** 21
    s.set_masses(1.0)
    s.positions = np.asarray(s.positions) + rng.normal(0, 0.1 * a, size=s.positions.shape)
    s.velocities = rng.normal(0, np.sqrt(0.001), size=s.positions.shape)
    s.set_pair_potential("(1-r)**2", r_cut=1.0)
    s.set_neighbor_list(skin=0.3, max_number_of_neighbors=64)
    s.set_integrator(time_step=0.005, target_temperature=0.001, temperature_damping_time=1.0)
    _check_against_oracle(s, "harm2", 1.0, 1.3)
    s.run(50)
    assert np.all(np.isfinite(np.asarray(s.positions)))
    _check_against_oracle(s, "harm2", 1.0, 1.3)


def test_2d_inverse_power_law_4m():
    from dompap_amd import Simulation
    rng = np.random.default_rng(5)
    s = Simulation(seed=55, precision="mixed")
    s.set_positions(unit_cell_coordinates=([0, 0],), cells=(2048, 2048), lattice_constants=(1.0, 1.0))
    assert s.number_of_particles() == 2 ** 22
    s.set_masses(1.0)
    s.positions = np.asarray(s.positions) + rng.normal(0, 0.05, size=s.positions.shape)
    s.velocities = rng.normal(0, 1.0, size=s.positions.shape)
    s.set_pair_potential("r**-12", r_cut=1.5)
    s.set_neighbor_list(skin=0.3, max_number_of_neighbors=32)
    s.set_integrator(time_step=0.002, target_temperature=1.0, temperature_damping_time=1.0)
    _check_against_oracle(s, ("ipl", 12), 1.5, 1.8)
    s.run(50)
    assert np.all(np.isfinite(np.asarray(s.positions)))
    _check_against_oracle(s, ("ipl", 12), 1.5, 1.8)
