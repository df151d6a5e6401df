"""A tile whose staged halo outgrows 16-bit neighbour entries.

The planner chooses the entry width from the mean occupancy; a dense region
in a dilute box (2-D, (1-r)**2, cutoff 1: 11 x 11-cell tiles with 13 x 13-cell
halos, one halo holding ~67,600 particles) overflows 16-bit entries at the
first build.  The build must switch to 32-bit entries by itself and give the
reference's lists (neighbor_list.py:7-39) and forces (potential.py:163-188)."""
import numpy as np
import pytest

from oracle import c_oracle as C

pytestmark = pytest.mark.gpu


def test_halo_beyond_16_bit_entries_rebuilds_with_32_bit():
    from dompap_amd.engine import Engine
    from dompap_amd.potential import device_potential
    rng = np.random.default_rng(17)
    L = 200.0
    dense = rng.uniform(10.0, 23.0, size=(13 * 13 * 400, 2))   # one tile's halo, 400 per cell
    dilute = rng.uniform(0.0, L, size=(3000, 2))
    x = np.ascontiguousarray(np.concatenate([dense, dilute]))
    n = x.shape[0]
    box = np.array([L, L])
    K = 1600
    eng = Engine(2, n, 1, "fp64")
    eng.set_box(box)
    dp = device_potential("(1-r)**2", 1.0)
    eng.load_potential(dp.code, 1.0, dp.u_rc, [[1.0]], [[1.0]])
    eng.set_neighbor_params(1.0, 0.0, K, 1)
    eng.set_field("positions", x)
    plan0 = eng.plan()
    assert plan0["entry_bits"] == 16 and plan0["halo_cells_per_tile"] == 169
    eng.build_neighbor_list()   # overflows 16-bit entries, rebuilds with 32-bit ones
    assert eng.plan()["entry_bits"] == 32
    nlo, cnt = C.neighbor_list(x, box, 1.0, K, "cells")
    assert np.array_equal(eng.neighbor_counts(), cnt)
    F = eng.compute(forces=True)["forces"]
    Fo = C.pair(x, box, nlo, "harm2", 1.0)["forces"]
    frms = float(np.sqrt(np.mean(np.sum(Fo ** 2, axis=1))))
    assert np.abs(F - Fo).max() / frms < 1e-10
    # rows as sets on a sample of dense particles
    nl = eng.neighbor_list()
    for i in rng.choice(dense.shape[0], 200, replace=False):
        assert set(nl[i][nl[i] >= 0].tolist()) == set(nlo[i][nlo[i] >= 0].tolist())
    eng.close()
