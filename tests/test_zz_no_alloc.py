"""The allocation-free entry point dp_rdf_counts_noalloc against
dp_rdf_counts (positions.py:58-87's pair counts): bit-exact counts on two
calls in a row, the second with a smaller input (fewer particles, fewer
bins), on both RDF kernels (cell tiles of the current list, all pairs), and
its size and state errors.  Calls go through the C-ABI (Engine)."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu
FCC = [(0, 0, 0), (.5, .5, 0), (.5, 0, .5), (0, .5, .5)]


def _lj_sim(cells, seed):
    from dompap_amd import Simulation
    s = Simulation(seed=seed)
    s.set_positions(unit_cell_coordinates=FCC, cells=(cells,) * 3)
    s.set_density(0.8442)
    s.set_masses(1.0)
    rng = np.random.default_rng(seed)
    s.positions = np.asarray(s.positions) + rng.normal(0, 0.08, size=np.asarray(s.positions).shape)
    s.set_pair_potential("4*(r**-12-r**-6)", r_cut=2.5)
    s.set_neighbor_list(skin=0.5, max_number_of_neighbors=256)
    s.get_forces()  # builds the list (the cell plan is then current)
    return s


def test_noalloc_rdf_matches_allocating_entry_point_twice():
    from dompap_amd import _lib
    big = _lib.load()  # noqa: F841  (the library is the native path under test)
    s1 = _lj_sim(10, 1)   # N = 4000
    s2 = _lj_sim(6, 2)    # N = 864: the second, smaller input
    e1, e2 = s1._engine, s2._engine
    bins1 = np.linspace(0.0, 2.9, 301)   # cell tiles (below the cell width)
    bins2 = np.linspace(0.5, 2.5, 51)
    for eng, bins, path in ((e1, bins1, 1), (e2, bins2, 1)):
        ref = eng.rdf_counts(bins)
        assert eng.rdf_path() == path
        got = eng.rdf_counts_noalloc(bins)
        assert eng.rdf_path() == path
        assert np.array_equal(got, ref)
        assert int(got.sum()) > 0
    # all-pairs kernel (last edge beyond the cell width), largest and smaller bin counts
    for eng, bins in ((e1, np.linspace(0.0, 9.0, 4097)), (e2, np.linspace(0.0, 7.0, 65))):
        ref = eng.rdf_counts(bins)
        assert eng.rdf_path() == 0
        got = eng.rdf_counts_noalloc(bins)
        assert eng.rdf_path() == 0
        assert np.array_equal(got, ref)
    # the scratch keeps the last call's counts: a repeated call starts from zero
    again = e2.rdf_counts_noalloc(np.linspace(0.0, 7.0, 65))
    assert np.array_equal(again, e2.rdf_counts(np.linspace(0.0, 7.0, 65)))


def test_noalloc_rdf_rejects_too_many_bins_and_stale_plan():
    s = _lj_sim(4, 3)
    eng = s._engine
    with pytest.raises(ValueError, match="at most 4097 bin edges"):
        eng.rdf_counts_noalloc(np.linspace(0.0, 2.0, 4098))
    # a changed box makes the cell plan stale until the list is rebuilt
    eng.set_box(np.asarray(s.box_vectors, dtype=np.float64) * 1.01)
    with pytest.raises(RuntimeError, match="cell plan is not current"):
        eng.rdf_counts_noalloc(np.linspace(0.0, 2.0, 11))
    eng.build_neighbor_list()
    got = eng.rdf_counts_noalloc(np.linspace(0.0, 2.0, 11))
    assert np.array_equal(got, eng.rdf_counts(np.linspace(0.0, 2.0, 11)))
