"""The measured workloads as parity cases, against the oracle:

* 3-D LJ, N = 4,096 on a 16^3 simple-cubic lattice at density 0.8, T = 0.8,
  r_cut = 2.5, skin 0.5, dt = 0.005, tau = 10, FP64 (the reference's own
  CPU-runnable case): the neighbour list bit-exact against the double loop
  (neighbor_list.py:7-39) with the reference's default K = 512, forces /
  energy / virial within 1e-10, a 60-step G-JF trajectory with injected
  noise against the oracle's Simulation.step restatement, and 1000 steps with
  device noise;
* 3-D LJ, N = 1,048,576 (FCC 64^3, density 0.8442): sorted neighbour rows
  equal to the C oracle's in both precision modes, at the first build and at
  a rebuild during steps (the mixed mode's packed 16-bit screen and guard
  bands decide every candidate there);
* 256 batched replicas, N log-uniform in [500, 20000], density uniform in
  [0.7, 1.0], T uniform in [0.5, 2.0], Gaussian-perturbed lattice starts,
  mixed precision: after steps, every replica's forces / energy / virial
  against the C oracle (FP64) on its own positions.

Tolerances: lists bit-exact; FP64 1e-10 relative; mixed 1e-5 of F_rms and
1e-6 relative for energy / virial (the replicas: 1e-6 and 1e-5 of
max(|value|, N))."""
import math

import numpy as np
import pytest

from oracle import c_oracle as C
from oracle import ref_numpy as R

pytestmark = pytest.mark.gpu
ONE = lambda n, m: 1.0  # noqa: E731
FCC = [(0, 0, 0), (.5, .5, 0), (.5, 0, .5), (0, .5, .5)]
LJ = "4*(r**-12-r**-6)"


def _frms(F):
    return float(np.sqrt(np.mean(np.sum(F ** 2, axis=1))))


def _sorted_rows(nl):
    big = np.iinfo(np.int32).max
    k = np.where(nl < 0, big, nl)
    k.sort(axis=1)
    return k


def _lj_4k(precision, seed=2024):
    from dompap_amd import Simulation
    s = Simulation(seed=seed, precision=precision)
    a = (1.0 / 0.8) ** (1.0 / 3.0)
    s.set_positions(unit_cell_coordinates=([0.0, 0.0, 0.0],), cells=(16, 16, 16), lattice_constants=(a, a, a))
    s.set_masses(1.0)
    rng = np.random.default_rng(seed)
    v = rng.normal(0.0, math.sqrt(0.8), size=(4096, 3))
    v -= v.mean(axis=0)
    s.velocities = v
    s.set_pair_potential(LJ, r_cut=2.5)
    s.set_neighbor_list(skin=0.5, max_number_of_neighbors=512)
    s.set_integrator(time_step=0.005, target_temperature=0.8, temperature_damping_time=10.0)
    return s, v


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_lj_4096_simple_cubic_lists_and_observables(precision):
    s, v = _lj_4k(precision)
    x0 = np.asarray(s.positions).copy()
    box = np.asarray(s.box_vectors)
    nl = R.neighbor_list_double_loop(x0, box, 3.0, 512)
    assert np.array_equal(s.neighbor_list, nl)
    # a perfect simple-cubic lattice: F = 0, and 7 shells of 6, 12, 8, 6, 24, 24, 0 inside the cut
    assert np.all(R.number_of_neighbors(nl) == 6 + 12 + 8 + 6 + 24 + 24)
    assert np.abs(s.get_forces()).max() < (1e-12 if precision == "fp64" else 1e-6)
    # thermalised: 40 steps with device noise, then the list of the last
    # rebuild and the observables at the current positions
    s.run(40)
    xr = np.asarray(s.positions_neighbour_list)
    assert np.array_equal(s.neighbor_list, R.neighbor_list_double_loop(xr, box, 3.0, 512))
    x = np.asarray(s.positions)
    nlx = R.neighbor_list_double_loop(x, box, 3.0, 512)
    u, f, c = R.make_pair_potential(LJ, 2.5)
    Fo = R.forces(x, box, f, nlx, ONE, ONE)
    Eo = R.total_energy(x, box, u, nlx, ONE, ONE)
    Wo = R.virial_double_loop(x, box, f, ONE, ONE)
    tol_f, tol_e = (1e-10, 1e-10) if precision == "fp64" else (1e-5, 1e-6)
    assert np.abs(s.get_forces() - Fo).max() / _frms(Fo) < tol_f
    assert abs(s.get_potential_energy() - Eo) < tol_e * abs(Eo)
    assert abs(s.get_virial() - Wo) < tol_e * abs(Wo)


def test_lj_4096_trajectory_with_injected_noise():
    s, v = _lj_4k("fp64", seed=7)
    x = np.asarray(s.positions).copy()
    box = np.asarray(s.box_vectors)
    o = R.OracleSim(x, box, v, np.ones(len(x)), LJ, 2.5, 0.5, 512, dt=0.005, T=0.8, tau=10.0)
    s.get_forces()  # the device's first build (at the start positions, like the oracle's)
    upd0 = s.number_of_neighbor_list_updates
    rng = np.random.default_rng(60)
    for k in range(60):
        xi = rng.normal(size=x.shape)
        s.step(noise=xi)
        o.step(xi)
    # the skin trigger fired on the same steps: as many rebuilds on both sides
    assert o.updates >= 2
    assert s.number_of_neighbor_list_updates - upd0 == o.updates - 1
    assert np.abs(np.asarray(s.positions) - o.x).max() < 1e-10
    assert np.abs(np.asarray(s.velocities) - o.v).max() < 1e-9
    assert np.array_equal(np.asarray(s.image_positions), o.img)


def test_lj_4096_thousand_steps():
    s, v = _lj_4k("fp64", seed=11)
    s.run(1000)
    x = np.asarray(s.positions)
    assert np.all(np.isfinite(x)) and np.all(np.isfinite(np.asarray(s.velocities)))
    assert s.number_of_steps == 1000
    box = np.asarray(s.box_vectors)
    assert np.all((x >= 0) & (x < box))
    # the simple-cubic start melts into a liquid (weak thermostat, tau = 10)
    assert 0.3 < s.get_temperature() < 3.0
    u = s.get_potential_energy() / 4096
    assert -6.5 < u < -2.0
    assert s.number_of_neighbor_list_updates > 10


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_lj_million_neighbor_sets_against_c_oracle(precision):
    from dompap_amd.engine import Engine
    from dompap_amd.potential import device_potential
    cells, rho = 64, 0.8442
    a = (4 / rho) ** (1 / 3)
    box = np.array([cells * a] * 3)
    rng = np.random.default_rng(31)
    x = (R.generate_positions(FCC, (cells,) * 3, (a, a, a)) + rng.normal(0, 0.07, (4 * cells ** 3, 3))) % box
    n = x.shape[0]
    K = 192
    eng = Engine(3, n, 1, precision)
    eng.set_box(box)
    dp = device_potential(LJ, 2.5)
    eng.load_potential(dp.code, 2.5, dp.u_rc, [[1.0]], [[1.0]])
    eng.set_neighbor_params(3.0, 0.5, K)
    eng.set_integrator(0.005, 0.72, 10.0)
    eng.set_seed(5, 0)
    eng.set_field("positions", x)
    eng.set_field("velocities", rng.normal(0, math.sqrt(0.72), (n, 3)))
    eng.build_neighbor_list()
    nlo, _ = C.neighbor_list(x, box, 3.0, K, "cells")
    assert np.array_equal(_sorted_rows(eng.neighbor_list()), _sorted_rows(nlo))
    del nlo
    upd0 = eng.stats()[0][0]
    eng.step(60, "gjf")
    assert eng.stats()[0][0] > upd0  # the list was rebuilt during the steps
    xr = eng.get_field("positions_neighbour_list")
    nlo, _ = C.neighbor_list(xr, box, 3.0, K, "cells")
    assert np.array_equal(_sorted_rows(eng.neighbor_list()), _sorted_rows(nlo))
    eng.close()


def replica_specs(n_rep=256, seed=3):
    """N log-uniform in [500, 20000], density in [0.7, 1.0], T in [0.5, 2.0]."""
    rng = np.random.default_rng(seed)
    n = np.rint(np.exp(rng.uniform(math.log(500), math.log(20000), n_rep))).astype(int)
    rho = rng.uniform(0.7, 1.0, n_rep)
    T = rng.uniform(0.5, 2.0, n_rep)
    return n, rho, T


def replica_start(n, rho, T, seed):
    """FCC lattice of 4 c^3 >= n sites filling a cubic box at density rho (a
    seeded subset of n sites), Gaussian displacements of 0.05, Gaussian
    velocities at T (zero momentum).  (A simple-cubic subset would put
    neighbours as close as 0.89 at density 1, where dt = 0.005 blows up.)"""
    rng = np.random.default_rng(seed)
    c = int(math.ceil((n / 4.0) ** (1.0 / 3.0) - 1e-9))
    L = (n / rho) ** (1.0 / 3.0)
    a = L / c
    sites = R.generate_positions(np.array(FCC, dtype=np.float64), (c,) * 3, (a,) * 3)
    x = sites[np.sort(rng.choice(sites.shape[0], size=n, replace=False))]
    box = np.full(3, L)
    x = (x + rng.normal(0, 0.05, x.shape)) % box
    v = rng.normal(0, math.sqrt(T), x.shape)
    v -= v.mean(axis=0)
    return x, v, box


def test_256_replicas_against_c_oracle():
    from dompap_amd import ReplicaBatch, Simulation
    n, rho, T = replica_specs()
    sims = []
    for r in range(len(n)):
        x, v, box = replica_start(int(n[r]), rho[r], T[r], 1000 + r)
        s = Simulation(seed=1000 + r, precision="mixed")
        s.positions = x
        s.box_vectors = box
        s.velocities = v
        s.betas = np.zeros_like(x)
        s.image_positions = np.zeros(x.shape, dtype=np.int32)
        s.masses = np.ones((x.shape[0], 1))
        s.particle_types = np.zeros((x.shape[0], 1), dtype=np.int32)
        s.set_pair_potential(LJ, r_cut=2.5)
        s.set_neighbor_list(skin=0.5, max_number_of_neighbors=256)
        s.set_integrator(time_step=0.005, target_temperature=float(T[r]), temperature_damping_time=1.0)
        sims.append(s)
    b = ReplicaBatch(sims)
    b.run(100)
    xs = b.get_field("positions")
    Fb = b.get_forces()
    Eb = b.get_potential_energy()
    Wb = b.get_virial()
    upd = b.number_of_neighbor_list_updates()
    assert np.all(upd >= 1)
    for r in range(len(n)):
        x = xs[r]
        box = np.asarray(sims[r].box_vectors, dtype=np.float64)
        assert np.all(np.isfinite(x))
        nl, _ = C.neighbor_list(x, box, 3.0, 256, "cells")
        o = C.pair(x, box, nl, "lj", 2.5, energy=True, virial=True)
        assert np.abs(Fb[r] - o["forces"]).max() / _frms(o["forces"]) < 1e-5, r
        # energy and virial: relative to max(|value|, N).  The mixed mode's
        # fixed-point coordinates (2^-26 units here: one scale for the batch,
        # set by the smallest boxes, which hold a single cell) perturb each
        # pair term; the virial's terms are large and cancel, so its error is
        # a random sum over pairs of ~1e-6 of |W| at a few hundred particles
        nr = x.shape[0]
        assert abs(Eb[r] - o["energy"]) < 1e-6 * max(abs(o["energy"]), nr), r
        assert abs(Wb[r] - o["virial"]) < 1e-5 * max(abs(o["virial"]), nr), r
    b.close()


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_single_cell_box_split_into_slices(precision):
    """A box narrower than three list cutoffs holds one cell; its 864
    particles are split into slices of home particles, one workgroup each.
    Lists (neighbor_list.py:7-39), forces, energy and virial as the oracle's,
    before and after steps."""
    from dompap_amd import Simulation
    s = Simulation(seed=12, precision=precision)
    s.set_positions(unit_cell_coordinates=FCC, cells=(6, 6, 6))
    s.set_density(1.2)  # L = 8.96 < 3 x 3.0
    s.set_masses(1.0)
    rng = np.random.default_rng(12)
    s.positions = np.asarray(s.positions) + rng.normal(0, 0.04, (864, 3))
    s.velocities = rng.normal(0, 1.0, (864, 3))
    s.set_pair_potential(LJ, r_cut=2.5)
    s.set_neighbor_list(skin=0.5, max_number_of_neighbors=400)
    s.set_integrator(time_step=0.002, target_temperature=1.0, temperature_damping_time=1.0)
    plan = s.get_kernel_plan()
    assert plan["cells"] == 1 and plan["tiles"] > 1, plan
    u, f, c = R.make_pair_potential(LJ, 2.5)
    box = np.asarray(s.box_vectors)
    for k in range(2):
        if k:
            s.run(30)
        x = np.asarray(s.positions)
        xr = np.asarray(s.positions_neighbour_list)
        assert np.array_equal(s.neighbor_list, R.neighbor_list_double_loop(xr, box, 3.0, 400))
        nl = R.neighbor_list_double_loop(x, box, 3.0, 400)
        Fo = R.forces(x, box, f, nl, ONE, ONE)
        Eo = R.total_energy(x, box, u, nl, ONE, ONE)
        Wo = R.virial_double_loop(x, box, f, ONE, ONE)
        tol_f, tol_e = (1e-10, 1e-10) if precision == "fp64" else (1e-5, 1e-6)
        assert np.abs(s.get_forces() - Fo).max() / _frms(Fo) < tol_f
        assert abs(s.get_potential_energy() - Eo) < tol_e * abs(Eo)
        assert abs(s.get_virial() - Wo) < tol_e * abs(Wo)
