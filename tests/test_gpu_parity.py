"""GPU parity: the HIP path (through the C-ABI, via dompap_amd) against the
oracle on the same seeded inputs, the reference's KATs/doctests/fixtures, and
size-independent properties at full size.

Tolerances: neighbour sets bit-exact; FP64 mode <= 1e-10 relative
(forces normalised by F_rms, energies/virials relative); mixed mode
<= 1e-5 (max |dF| / F_rms) and 1e-6 relative for energy / virial."""
import math
import os

import numpy as np
import pytest

from oracle import c_oracle as C
from oracle import ref_numpy as R

pytestmark = pytest.mark.gpu
ONE = lambda n, m: 1.0  # noqa: E731
FCC = [(0, 0, 0), (.5, .5, 0), (.5, 0, .5), (0, .5, .5)]


def _sim(precision="fp64", **kw):
    from dompap_amd import Simulation
    return Simulation(seed=7, precision=precision, **kw)


def _lj_system(precision, cells=6, density=0.9, noise=0.05, skin=0.5, seed=0, K=256):
    s = _sim(precision)
    s.set_positions(unit_cell_coordinates=FCC, cells=(cells,) * 3)
    s.set_density(density)
    s.set_masses(1.0)
    rng = np.random.default_rng(seed)
    x = np.asarray(s.positions) + rng.normal(0, noise, size=s.positions.shape)
    s.positions = x
    s.set_pair_potential("4*(r**-12-r**-6)", r_cut=2.5)
    s.set_neighbor_list(skin=skin, max_number_of_neighbors=K)
    return s, x


def _frms(F):
    return float(np.sqrt(np.mean(np.sum(F ** 2, axis=1))))


# ------------------------------------------------------------ doctests / KATs
def test_doctest_force_on_particle_0():
    # Simulation.py:600-605
    s = _sim()
    s.positions[0] = [0.5, 0.0, 0.0]
    F = s.get_forces()
    assert np.array_equal(F[0], [-1.0, 0.0, 0.0])


def test_doctest_energy_scaled_box():
    # Simulation.py:551-555 (167.06442443573454; last digits are summation order)
    s = _sim()
    s.scale_box(0.5)
    E = s.get_potential_energy()
    assert abs(E - 167.06442443573454) / 167.06442443573454 < 1e-13


def test_doctest_laplacian_virial_pressure_neighbours():
    # Simulation.py:1211-1215, :952-955, :976-979, :327-333
    s = _sim()
    assert s.get_laplacian()[0] == 12.0
    assert s.get_virial() == 0.0
    s.set_random_velocities(temperature=0.0)
    assert s.get_pressure() == 0.0
    from dompap_amd.neighbor_list import get_number_of_neighbors
    assert np.max(get_number_of_neighbors(s.neighbor_list)) == 26
    assert s.number_of_neighbor_list_updates >= 3


def test_doctest_time_after_ten_steps():
    # Simulation.py:938-942
    s = _sim()
    s.set_integrator(time_step=0.01, target_temperature=1.0, temperature_damping_time=0.1)
    for _ in range(10):
        s.step()
    assert s.get_time() == 0.1


def test_doctest_wrap_into_box():
    # Simulation.py:662-675
    s = _sim()
    s.positions[0] = [6.0, 0.0, 0.0]
    s.wrap_into_box()
    assert np.array_equal(s.positions[0], [1.0, 0.0, 0.0])
    assert np.array_equal(s.image_positions[0], [1, 0, 0])


def test_kat_energy_force_sigma2_eps4():
    # tests/test_potential.py:81-140 through the Simulation API
    s = _sim()
    s.positions = np.array([[0, 0, 0], [1, 0, 0]], float)
    s.box_vectors = np.array([3.0, 3, 3])
    s.set_masses(1.0)
    s.set_pair_potential("(1-r)**2", r_cut=1.0)
    s.set_pair_potential_parameters(sigma=2.0, epsilon=4.0)
    assert s.get_potential_energy() == 1.0
    assert np.allclose(s.get_forces(), [[-4, 0, 0], [4, 0, 0]])


def test_fixture_harmonic_repulsive_3d(golden_dir):
    # examples/harmonic_repulsive_3d.{csv,toml}
    from dompap_amd import Simulation
    s = Simulation(seed=1)
    meta = s.from_disk(os.path.join(golden_dir, "harmonic_repulsive_3d.csv"),
                       os.path.join(golden_dir, "harmonic_repulsive_3d.toml"))
    assert abs(s.get_virial() - meta["virial"]) / meta["virial"] < 1e-12
    assert s.get_kinetic_energy() == meta["kinetic_energy"]
    assert s.get_temperature() == meta["temperature"]
    assert abs(s.get_pressure() - meta["pressure"]) < 1e-14
    assert abs(s.get_potential_energy() - 0.31658024095731246) < 1e-13  # brute force, not the TOML value
    assert s.number_of_neighbor_list_updates >= meta["number_of_neighbor_list_updates"]


# ------------------------------------------------------------ parity vs oracle
@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_lj_parity_against_oracle(precision):
    s, x = _lj_system(precision)
    box = np.asarray(s.box_vectors)
    F = s.get_forces()
    E = s.get_potential_energy()
    W = s.get_virial()
    L = s.get_laplacian()
    nl = s.neighbor_list
    u, f, c = R.make_pair_potential("4*(r**-12-r**-6)", 2.5)
    nlo = R.neighbor_list_double_loop(x, box, 3.0, 256)
    assert np.array_equal(nl, nlo)  # bit-exact neighbour sets and row order
    Fo = R.forces(x, box, f, nlo, ONE, ONE)
    Eo = R.total_energy(x, box, u, nlo, ONE, ONE)
    Wo = R.virial_double_loop(x, box, f, ONE, ONE)
    Lo = R.laplacian(x, box, f, c, nlo, ONE, ONE)
    tol_f, tol_e = (1e-10, 1e-10) if precision == "fp64" else (1e-5, 1e-6)
    assert np.abs(F - Fo).max() / _frms(Fo) < tol_f
    assert abs(E - Eo) / abs(Eo) < tol_e
    assert abs(W - Wo) / abs(Wo) < tol_e
    assert np.abs(L - Lo).max() / np.abs(Lo).max() < (1e-10 if precision == "fp64" else 1e-5)


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_gjf_step_with_injected_noise(precision):
    s, x = _lj_system(precision)
    box = np.asarray(s.box_vectors)
    rng = np.random.default_rng(1)
    v = rng.normal(0, 0.8, size=x.shape)
    s.velocities = v
    s.set_integrator(time_step=0.005, target_temperature=0.8, temperature_damping_time=10.0)
    F = s.get_forces()
    xi = rng.normal(size=x.shape)
    s.step(noise=xi)
    m = np.ones((x.shape[0], 1))
    xo, vo, bo = R.make_one_step(x, v, np.zeros_like(x), F, m, 0.005, 0.8, 10.0, xi)
    img = np.zeros(x.shape, np.int32)
    R.wrap_into_box(xo, img, box)
    # forces given, the update is the reference's arithmetic: identical bits
    assert np.array_equal(np.asarray(s.betas), bo)
    assert np.array_equal(np.asarray(s.velocities), vo)
    assert np.array_equal(np.asarray(s.positions), xo)
    assert np.array_equal(np.asarray(s.image_positions), img)


def test_integrator_free_functions_bit_exact():
    from dompap_amd.integrator import make_one_step, make_one_step_leap_frog
    rng = np.random.default_rng(2)
    x, v, b, F = rng.normal(size=(4, 50, 3))
    m = rng.uniform(0.5, 2.0, (50, 1))
    xi = rng.normal(size=x.shape)
    got = make_one_step(x, v, b, F, m, 0.003, 1.3, 0.7, noise=xi)
    exp = R.make_one_step(x, v, b, F, m, 0.003, 1.3, 0.7, xi)
    for g, e in zip(got, exp):
        assert np.array_equal(g, e)
    got = make_one_step_leap_frog(x, v, F, m, 0.003)
    exp = R.make_one_step_leap_frog(x, v, F, m, 0.003)
    for g, e in zip(got, exp):
        assert np.array_equal(g, e)


def test_device_noise_is_philox_box_muller():
    # beta' = sqrt(2 alpha T dt) * xi with zero force and velocity
    from dompap_amd.engine import Engine
    from dompap_amd.potential import device_potential
    n = 1000
    eng = Engine(3, n, 1, "fp64")
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
    eng.step(1)
    beta = eng.get_field("betas")
    xi = C.normals(n, 3, 12345, 77)
    expect = math.sqrt(2 * (1.0 / tau) * T * dt) * xi
    assert np.abs(beta - expect).max() < 1e-13


def test_neighbor_list_free_functions_kats():
    from dompap_amd.neighbor_list import (get_neighbor_list_cell_list, get_neighbor_list_cell_list_3d,
                                          get_neighbor_list_double_loop, neighbor_list_is_old)
    x = np.array([[0, 0, 0], [1, 0, 0], [0, 1, 0], [0, 0, 1], [2, 2, 2], [5, 5, 0]], float)
    box = np.array([6.0, 6, 6])
    assert np.array_equal(get_neighbor_list_double_loop(x, box, 1.5, 4),
                          [[1, 2, 3, 5], [0, 2, 3, -1], [0, 1, 3, -1], [0, 1, 2, -1], [-1] * 4, [0, -1, -1, -1]])
    assert np.array_equal(get_neighbor_list_cell_list_3d(x, box, 1.5, 4),
                          [[5, 1, 2, 3], [0, 2, 3, -1], [0, 1, 3, -1], [0, 1, 2, -1], [-1] * 4, [0, -1, -1, -1]])
    x2 = np.array([[0, 0], [1, 0], [0, 1], [2, 2], [5, 5]], float)
    assert np.array_equal(get_neighbor_list_double_loop(x2, np.array([6.0, 6]), 1.5, 4),
                          [[1, 2, 4, -1], [0, 2, -1, -1], [0, 1, -1, -1], [-1] * 4, [0, -1, -1, -1]])
    assert np.array_equal(get_neighbor_list_cell_list(x2, np.array([6.0, 6]), 1.5, 4),
                          [[4, 1, 2, -1], [0, 2, -1, -1], [0, 1, -1, -1], [-1] * 4, [0, -1, -1, -1]])
    pos = np.zeros((2, 3))
    old = np.zeros((2, 3))
    b5 = np.array([5.0, 5, 5])
    assert not neighbor_list_is_old(pos, old, b5, 0.5)
    old[0] = [0.25 - 0.01, 0, 0]
    assert not neighbor_list_is_old(pos, old, b5, 0.5)
    old[0] = [0.25 + 0.01, 0, 0]
    assert neighbor_list_is_old(pos, old, b5, 0.5)
    old[0] = [5.0, 0, 0]
    assert not neighbor_list_is_old(pos, old, b5, 0.5)


@pytest.mark.parametrize("dim,shape,pot,r_cut,skin", [(2, (40, 40), "(1-r)**2", 1.0, 0.3),
                                                      (4, (6, 6, 6, 6), "(1-r)**2", 1.0, 0.3),
                                                      (2, (40, 40), "r**-12", 1.5, 0.3)])
@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_any_dimension_parity(dim, shape, pot, r_cut, skin, precision):
    from dompap_amd import Simulation
    rng = np.random.default_rng(dim)
    s = Simulation(seed=3, precision=precision)
    s.set_positions(unit_cell_coordinates=(np.zeros(dim),), cells=shape, lattice_constants=(0.92,) * dim)
    s.set_masses(1.0)
    x = np.asarray(s.positions) + rng.normal(0, 0.06, size=s.positions.shape)
    s.positions = x
    s.set_pair_potential(pot, r_cut=r_cut)
    s.set_neighbor_list(skin=skin, max_number_of_neighbors=128)
    box = np.asarray(s.box_vectors)
    F = s.get_forces()
    E = s.get_potential_energy()
    W = s.get_virial()
    nl = s.neighbor_list
    u, f, c = R.make_pair_potential(pot, r_cut)
    nlo = R.neighbor_list_double_loop(x, box, r_cut + skin, 128)
    assert np.array_equal(nl, nlo)
    Fo = R.forces(x, box, f, nlo, ONE, ONE)
    Eo = R.total_energy(x, box, u, nlo, ONE, ONE)
    Wo = R.virial_neighbor_list(x, box, f, nlo, ONE, ONE)
    tol = 1e-10 if precision == "fp64" else 1e-5
    assert np.abs(F - Fo).max() / _frms(Fo) < tol
    assert abs(E - Eo) / abs(Eo) < tol
    assert abs(W - Wo) / abs(Wo) < tol


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_binary_mixture_types(precision):
    # per-type sigma table (Wahnstrom-type: sigma_BB = 1/1.2, sigma_AB mean)
    from dompap_amd import Simulation
    rng = np.random.default_rng(4)
    s = Simulation(seed=5, precision=precision)
    s.set_positions(unit_cell_coordinates=([0, 0, 0],), cells=(10, 10, 10), lattice_constants=(1.1, 1.1, 1.1))
    n = s.number_of_particles()
    s.set_masses(1.0)
    types = np.zeros(n, dtype=np.int32)
    types[rng.permutation(n)[: n // 2]] = 1
    s.set_types(types.tolist())
    sig = np.array([[1.0, (1 + 1 / 1.2) / 2], [(1 + 1 / 1.2) / 2, 1 / 1.2]])
    x = np.asarray(s.positions) + rng.normal(0, 0.05, size=s.positions.shape)
    s.positions = x
    s.set_pair_potential("4*(r**-12-r**-6)", r_cut=2.5)
    s.set_pair_potential_parameters(sigma=lambda a, b: sig[types[a], types[b]], epsilon=1.0)
    s.set_neighbor_list(skin=0.5, max_number_of_neighbors=256)
    box = np.asarray(s.box_vectors)
    F = s.get_forces()
    E = s.get_potential_energy()
    sf = lambda a, b: sig[types[a], types[b]]  # noqa: E731
    u, f, c = R.make_pair_potential("4*(r**-12-r**-6)", 2.5)
    nlo = R.neighbor_list_double_loop(x, box, 1.0 * 2.5 + 0.5, 256)
    assert np.array_equal(s.neighbor_list, nlo)
    Fo = R.forces(x, box, f, nlo, sf, ONE)
    Eo = R.total_energy(x, box, u, nlo, sf, ONE)
    tol = 1e-10 if precision == "fp64" else 1e-5
    assert np.abs(F - Fo).max() / _frms(Fo) < tol
    assert abs(E - Eo) / abs(Eo) < tol


def test_overflow_raises_value_error():
    s, x = _lj_system("fp64", K=16)
    with pytest.raises(ValueError):
        s.get_forces()


def test_graph_replay_matches_direct_steps():
    # run(k) replays captured hipGraph blocks; step() launches directly
    a, x = _lj_system("mixed", cells=5)
    b, _ = _lj_system("mixed", cells=5)
    v = np.random.default_rng(3).normal(0, 1.0, size=x.shape)
    for sim in (a, b):
        sim.velocities = v
        sim.set_integrator(time_step=0.005, target_temperature=1.0, temperature_damping_time=1.0)
    a.run(130)
    for _ in range(130):
        b.step()
    assert np.array_equal(np.asarray(a.positions), np.asarray(b.positions))
    assert np.array_equal(np.asarray(a.velocities), np.asarray(b.velocities))
    assert a.number_of_neighbor_list_updates == b.number_of_neighbor_list_updates


def test_step_blocks_of_any_length_match_direct_steps():
    # step blocks stop at the first skin trigger and at the requested count;
    # odd block lengths change the buffer parity between blocks
    for block, n in ((7, 53), (16, 9), (64, 100)):
        a, x = _lj_system("fp64", cells=5)
        b, _ = _lj_system("fp64", cells=5)
        v = np.random.default_rng(block).normal(0, 1.5, size=x.shape)
        for sim in (a, b):
            sim.velocities = v
            sim.set_integrator(time_step=0.005, target_temperature=1.5, temperature_damping_time=1.0)
        a._device().set_graph_block(block)
        a.run(n)
        for _ in range(n):
            b.step()
        assert np.array_equal(np.asarray(a.positions), np.asarray(b.positions)), block
        assert np.array_equal(np.asarray(a.betas), np.asarray(b.betas)), block
        assert a.number_of_neighbor_list_updates == b.number_of_neighbor_list_updates, block
        assert a.number_of_steps == b.number_of_steps == n


def test_trajectory_parity_fp64_short():
    # 20 G-JF steps with the device noise stream replayed by the oracle
    s, x = _lj_system("fp64", cells=4, density=0.8)
    box = np.asarray(s.box_vectors)
    s.set_integrator(time_step=0.005, target_temperature=0.8, temperature_damping_time=10.0)
    rng = np.random.default_rng(9)
    v = rng.normal(0, 0.9, size=x.shape)
    s.velocities = v
    o = R.OracleSim(x, box, v, np.ones(len(x)), "4*(r**-12-r**-6)", 2.5, 0.5, 256, dt=0.005, T=0.8, tau=10.0)
    for k in range(20):
        xi = rng.normal(size=x.shape)
        s.step(noise=xi)
        o.step(xi)
    assert np.abs(np.asarray(s.positions) - o.x).max() < 1e-10
    assert np.abs(np.asarray(s.velocities) - o.v).max() < 1e-9


# ------------------------------------------------------------ full-size properties
@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_perfect_fcc_million(precision):
    # N = 1,048,576 perfect FCC: F == 0 by symmetry, U/N = closed-form lattice sum
    from dompap_amd.engine import Engine
    from dompap_amd.potential import device_potential
    cells, rho = 64, 0.8442
    a = (4 / rho) ** (1 / 3)
    x = R.generate_positions(FCC, (cells,) * 3, (a, a, a))
    n = x.shape[0]
    eng = Engine(3, n, 1, precision)
    eng.set_box([cells * a] * 3)
    dp = device_potential("4*(r**-12-r**-6)", 2.5)
    eng.load_potential(dp.code, 2.5, dp.u_rc, [[1.0]], [[1.0]])
    eng.set_neighbor_params(3.0, 0.5, 256)
    eng.set_field("positions", x)
    eng.build_neighbor_list()
    out = eng.compute(forces=True, energy=True, virial=True)
    u, f, c = R.make_pair_potential("4*(r**-12-r**-6)", 2.5)
    z = {1: 12, 2: 6, 3: 24, 4: 12, 5: 24, 6: 8, 7: 48, 8: 6, 9: 36, 10: 24, 11: 24, 12: 24, 13: 72}
    e = w = 0.0
    kbar = 0
    for k, zk in z.items():
        r = math.sqrt(k / 2) * a
        if r <= 2.5:
            e += zk * float(u(r))
            w += zk * float(f(r)) * r
        if r < 3.0:
            kbar += zk
    tol = 1e-10 if precision == "fp64" else 2e-6
    assert abs(out["energy"][0] / n - e / 2) < tol * abs(e)
    assert abs(out["virial"][0] / n - w / 6) < tol * max(abs(w), 1.0)
    assert np.abs(out["forces"]).max() < (1e-10 if precision == "fp64" else 1e-3)
    counts = eng.neighbor_counts()
    assert np.all(counts == kbar)
    eng.close()


def test_million_random_parity_against_c_oracle():
    # N = 1,048,576 perturbed FCC in FP64 against the C oracle (cell search)
    from dompap_amd.engine import Engine
    from dompap_amd.potential import device_potential
    cells, rho = 64, 0.8442
    a = (4 / rho) ** (1 / 3)
    box = np.array([cells * a] * 3)
    rng = np.random.default_rng(11)
    x = (R.generate_positions(FCC, (cells,) * 3, (a, a, a)) + rng.normal(0, 0.07, (4 * cells ** 3, 3))) % box
    n = x.shape[0]
    res = {}
    for prec in ("fp64", "mixed"):
        eng = Engine(3, n, 1, prec)
        eng.set_box(box)
        dp = device_potential("4*(r**-12-r**-6)", 2.5)
        eng.load_potential(dp.code, 2.5, dp.u_rc, [[1.0]], [[1.0]])
        eng.set_neighbor_params(3.0, 0.5, 256)
        eng.set_field("positions", x)
        eng.build_neighbor_list()
        res[prec] = eng.compute(forces=True, energy=True, virial=True)
        res[prec]["counts"] = eng.neighbor_counts()
        eng.close()
    nl, cnt = C.neighbor_list(x, box, 3.0, 256, "cells")
    assert np.array_equal(res["fp64"]["counts"], cnt)
    o = C.pair(x, box, nl, "lj", 2.5, energy=True, virial=True)
    frms = _frms(o["forces"])
    assert np.abs(res["fp64"]["forces"] - o["forces"]).max() / frms < 1e-10
    assert abs(res["fp64"]["energy"][0] - o["energy"]) / abs(o["energy"]) < 1e-10
    assert abs(res["fp64"]["virial"][0] - o["virial"]) / abs(o["virial"]) < 1e-10
    assert np.abs(res["mixed"]["forces"] - o["forces"]).max() / frms < 1e-5
    assert abs(res["mixed"]["energy"][0] - o["energy"]) / abs(o["energy"]) < 1e-6
    # Newton's third law at full size
    assert np.abs(res["fp64"]["forces"].sum(0)).max() / frms < 1e-8


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_fused_and_split_step_kernels_agree(precision):
    """The fused force + integrate kernel and the force-then-integrate pair
    run the same arithmetic in the same lane order: identical trajectories."""
    runs = []
    for fused in (0, 1):
        s, x = _lj_system(precision, cells=6, density=0.9)
        s.set_masses(1.0)
        s.set_integrator(time_step=0.005, target_temperature=1.0, temperature_damping_time=1.0)
        s.set_kernel_tuning(2, 0, fused)
        s.velocities = np.random.default_rng(9).normal(0, 1, size=x.shape)
        s.betas = np.zeros_like(x)
        s.run(30)
        runs.append((np.asarray(s.positions).copy(), np.asarray(s.velocities).copy()))
    assert np.array_equal(runs[0][0], runs[1][0])
    assert np.array_equal(runs[0][1], runs[1][1])


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_inhomogeneous_tiles_beyond_lds_capacity(precision):
    """Half of the box empty: the planner sizes the LDS record areas from the
    mean occupancy, so the dense tiles take the global-memory paths of the
    emit and the pair kernels (the densest still do after the measured halo
    raised the areas); lists and forces must not change.  Two lanes per
    particle keep the tiles of several cells whose halos outgrow the areas
    (the planner would give this small system 16 lanes and one-cell tiles)."""
    from dompap_amd import Simulation
    rng = np.random.default_rng(31)
    g = np.arange(12)
    half = np.stack(np.meshgrid(g, np.arange(24), np.arange(24), indexing="ij"), -1).reshape(-1, 3) * 0.95
    x = half + rng.normal(0, 0.05, size=half.shape)
    box = np.array([24 * 0.95] * 3)
    x = x % box
    s = Simulation(seed=3, precision=precision)
    s.box_vectors = box
    s.positions = x
    s.velocities = np.zeros_like(x)
    s.betas = np.zeros_like(x)
    s.image_positions = np.zeros(x.shape, dtype=np.int32)
    s.set_masses(1.0)
    s.particle_types = np.zeros((x.shape[0], 1), dtype=np.int32)
    s.set_pair_potential("4*(r**-12-r**-6)", r_cut=2.5)
    s.set_neighbor_list(skin=0.5, max_number_of_neighbors=256)
    s.set_kernel_tuning(threads_per_particle=2)
    F = s.get_forces()
    plan = s.get_kernel_plan()
    nl = np.asarray(s.neighbor_list)
    nlo, _ = C.neighbor_list(x, box, 3.0, 256, "cells")
    assert np.array_equal(nl, nlo)
    Fo = C.pair(x, box, nlo, "lj", 2.5)["forces"]
    tol = 1e-10 if precision == "fp64" else 1e-5
    assert np.abs(F - Fo).max() / _frms(Fo) < tol
    # a tile inside the occupied half holds about twice the mean occupancy
    dense_halo = 2.0 * x.shape[0] / plan["cells"] * plan["halo_cells_per_tile"]
    assert dense_halo > plan["lds_records"]
    # the largest halo that the emit measured raises the LDS record areas
    # of the later launches above the mean-occupancy rule (as far as four
    # workgroups per compute unit allow); the densest tiles still exceed them
    h = x.shape[0] / plan["cells"] * plan["halo_cells_per_tile"]
    rule = math.ceil(min(1.5 * h, h + 3 * math.sqrt(h))) + 64
    assert plan["lds_records"] > rule and plan["lds_records_emit"] > rule
    F2 = s.get_forces()
    assert np.abs(F2 - Fo).max() / _frms(Fo) < tol


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_dense_cells_several_particle_groups_per_wavefront(precision):
    """Uniform random gas at density 6 with a list cutoff of 2: ~48 particles
    per cell and ~200 neighbours each, so every wavefront of the emit takes
    its particles of a home cell in more than one group; lists in the
    oracle's order and forces against the oracle."""
    from dompap_amd import Simulation
    rng = np.random.default_rng(41)
    L = 16.0
    n = 6 * 16 ** 3
    x = rng.uniform(0.0, L, size=(n, 3))
    box = np.array([L] * 3)
    s = Simulation(seed=5, precision=precision)
    s.box_vectors = box
    s.positions = x
    s.velocities = np.zeros_like(x)
    s.betas = np.zeros_like(x)
    s.image_positions = np.zeros(x.shape, dtype=np.int32)
    s.set_masses(1.0)
    s.particle_types = np.zeros((n, 1), dtype=np.int32)
    s.set_pair_potential("(1-r)**2", r_cut=1.0)
    s.set_neighbor_list(skin=1.0, max_number_of_neighbors=384)
    F = s.get_forces()
    nl = np.asarray(s.neighbor_list)
    nlo, cnt = C.neighbor_list(x, box, 2.0, 384, "cells")
    assert cnt.max() > 150
    assert np.array_equal(nl, nlo)
    Fo = C.pair(x, box, nlo, "harm2", 1.0)["forces"]
    tol = 1e-10 if precision == "fp64" else 1e-5
    assert np.abs(F - Fo).max() / _frms(Fo) < tol


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_dilute_3d_gas_lane_emit(precision):
    """3-D LJ gas at density 0.08: ~2 particles per cell, so the emit takes
    one lane per particle; lists in the oracle's order and forces against
    the oracle."""
    from dompap_amd import Simulation
    rng = np.random.default_rng(43)
    L = 40.0
    g = np.arange(17) * (L / 17)
    x = np.stack(np.meshgrid(g, g, g, indexing="ij"), -1).reshape(-1, 3)
    x = (x + rng.normal(0, 0.3, size=x.shape)) % L  # jittered sparse lattice: no overlaps
    n = x.shape[0]
    box = np.array([L] * 3)
    s = Simulation(seed=9, precision=precision)
    s.box_vectors = box
    s.positions = x
    s.velocities = np.zeros_like(x)
    s.betas = np.zeros_like(x)
    s.image_positions = np.zeros(x.shape, dtype=np.int32)
    s.set_masses(1.0)
    s.particle_types = np.zeros((n, 1), dtype=np.int32)
    s.set_pair_potential("4*(r**-12-r**-6)", r_cut=2.5)
    s.set_neighbor_list(skin=0.5, max_number_of_neighbors=64)
    F = s.get_forces()
    nl = np.asarray(s.neighbor_list)
    nlo, _ = C.neighbor_list(x, box, 3.0, 64, "cells")
    assert np.array_equal(nl, nlo)
    Fo = C.pair(x, box, nlo, "lj", 2.5)["forces"]
    tol = 1e-10 if precision == "fp64" else 1e-5
    assert np.abs(F - Fo).max() / _frms(Fo) < tol
