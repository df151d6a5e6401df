"""Radial distribution function on the GPU (dp_rdf_counts through
Simulation.get_radial_distribution_function) against the oracle's counts of
positions.py:58-87: bit-exact pair counts, hence bit-exact g(r), on both
kernels (cell tiles of the current list, all pairs), edges that fall exactly
on realised distances, repeated edges, more bins than one LDS histogram holds,
and 2-D / 4-D boxes.  At sizes the oracle cannot finish quickly the two
kernels are compared with each other."""
import numpy as np
import pytest

from oracle import ref_numpy as R

pytestmark = pytest.mark.gpu
FCC = [(0, 0, 0), (.5, .5, 0), (.5, 0, .5), (0, .5, .5)]


def _sim(dim_cells, precision="fp64", pot=("4*(r**-12-r**-6)", 2.5), skin=0.5, K=256):
    from dompap_amd import Simulation
    s = Simulation(seed=7, precision=precision)
    s.set_positions(unit_cell_coordinates=([0.0] * len(dim_cells),), cells=dim_cells,
                    lattice_constants=(1.0,) * len(dim_cells))
    s.set_masses(1.0)
    s.set_pair_potential(pot[0], r_cut=pot[1])
    s.set_neighbor_list(skin=skin, max_number_of_neighbors=K)
    return s


def _check(s, r_bins, path):
    from dompap_amd.positions import normalise_radial_histogram
    x = np.asarray(s.positions)
    box = np.asarray(s.box_vectors, dtype=np.float64)
    r, g = s.get_radial_distribution_function(r_bins)
    assert s._engine.rdf_path() == path
    c = R.rdf_counts(x, box, r_bins)
    r0, g0 = normalise_radial_histogram(c, np.asarray(r_bins, dtype=np.float64), x.shape[0] ** 2, box, x.shape[1])
    assert np.array_equal(s._engine.rdf_counts(r_bins).astype(np.int64), c)
    assert np.array_equal(r, r0) and np.array_equal(g, g0, equal_nan=True)
    return c


def test_rdf_random_gas_all_pairs():
    # tests/test_positions.py:66-76's sizes: 1000 points, box 10, 19 bins
    s = _sim((10, 10, 10))
    rng = np.random.default_rng(11)
    s.positions = rng.random((1000, 3)) * 10.0
    _check(s, np.linspace(0.1, 5, 20), path=0)
    # every pair in one bin: N^2 counts, self pairs included
    c = s._engine.rdf_counts([0.0, 100.0])
    assert int(c[0]) == 1000 * 1000


def test_rdf_lj_fluid_cell_tiles_and_all_pairs():
    for precision in ("fp64", "mixed"):
        from dompap_amd import Simulation
        s = Simulation(seed=3, precision=precision)
        s.set_positions(unit_cell_coordinates=FCC, cells=(8, 8, 8))
        s.set_density(0.8442)
        s.set_masses(1.0)
        s.set_pair_potential("4*(r**-12-r**-6)", r_cut=2.5)
        s.set_neighbor_list(skin=0.5, max_number_of_neighbors=256)
        s.set_integrator(time_step=0.005, target_temperature=1.2, temperature_damping_time=1.0)
        rng = np.random.default_rng(5)
        s.velocities = rng.normal(0, np.sqrt(1.2), size=np.asarray(s.positions).shape)
        s.run(40)
        # last edge + twice the displacement since the rebuild below the cell width
        _check(s, np.linspace(0.0, 2.5, 251), path=1)
        _check(s, np.linspace(0.5, 6.0, 56), path=0)


def test_rdf_exact_and_repeated_edges():
    # simple cubic, a = 1: distances 0, 1, sqrt2, sqrt3, 2 exactly
    s = _sim((8, 8, 8), pot=("(1-r)**2", 1.0), skin=0.3, K=64)
    n = 512
    c = _check(s, [0.0, 0.5, 1.0], path=1)  # the closed last bin holds r = 1
    assert c.tolist() == [n, 6 * n]
    c = _check(s, [0.0, 1.0, 1.0, np.sqrt(2.0), 2.0, 2.0], path=0)
    assert c.tolist() == [n, 0, 6 * n, 20 * n, 6 * n]
    c = _check(s, [-1.0, 0.0, 0.0, 1.0], path=1)
    assert c.tolist() == [0, 0, n + 6 * n]


def test_rdf_more_bins_than_one_histogram():
    s = _sim((10, 10, 10))
    rng = np.random.default_rng(12)
    s.positions = rng.random((1000, 3)) * 10.0
    _check(s, np.linspace(0.0, 5.0, 9001), path=0)
    # the list was built at these positions (3 cells of 3.33 per dimension)
    _check(s, np.linspace(0.2, 2.5, 6001), path=1)


def test_rdf_2d_and_4d():
    rng = np.random.default_rng(13)
    s = _sim((30, 50))
    s.positions = rng.random((1500, 2)) * np.array([30.0, 50.0])
    _check(s, np.linspace(0.0, 12.0, 97), path=0)
    s = _sim((5, 5, 6, 6), pot=("(1-r)**2", 1.0), skin=0.3, K=128)
    s.positions = np.asarray(s.positions) + rng.normal(0, 0.05, size=(900, 4))
    _check(s, np.linspace(0.0, 1.2, 61), path=1)
    _check(s, np.linspace(0.0, 3.0, 61), path=0)


def test_rdf_cell_tiles_match_all_pairs_at_scale():
    from dompap_amd import Simulation
    s = Simulation(seed=9, precision="mixed")
    s.set_positions(unit_cell_coordinates=FCC, cells=(20, 20, 20))
    s.set_density(0.8442)
    s.set_masses(1.0)
    s.set_pair_potential("4*(r**-12-r**-6)", r_cut=2.5)
    s.set_neighbor_list(skin=0.5, max_number_of_neighbors=256)
    s.set_integrator(time_step=0.005, target_temperature=0.8, temperature_damping_time=1.0)
    s.velocities = np.random.default_rng(3).normal(0, np.sqrt(0.8), size=np.asarray(s.positions).shape)
    s.run(60)
    e = s._engine
    edges = np.linspace(0.5, 2.5, 201)
    c_cells = e.rdf_counts(edges).astype(np.int64)
    assert e.rdf_path() == 1
    c_all = e.rdf_counts(np.append(edges, 1e3)).astype(np.int64)
    assert e.rdf_path() == 0
    # identical bins, except that r == 2.5 exactly sits in the closed last
    # bin of the first call and in the extra bin of the second
    assert np.array_equal(c_cells[:-1], c_all[:-2])
    assert c_cells[-1] >= c_all[-2] and c_cells[-1] - c_all[-2] <= c_all[-1]
    n = s.number_of_particles()
    assert int(e.rdf_counts([0.0, 0.5, 1e3]).sum()) == n * n
