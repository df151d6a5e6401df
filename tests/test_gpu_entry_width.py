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


def test_mixed_halo_beyond_byte_offset_entries_rebuilds_with_32_bit():
    """Mixed precision: tiles whose halo fits the pair kernels' LDS record
    area take byte offsets of their 8-byte staged records (2-D, one type:
    shift 3) as entries; the dense tile's halo does not, so it keeps plain
    staged-halo indices, which overflow 16 bits as in FP64 mode.  Its 11 x
    11-cell home box is too wide for a sentinel record, so rows are padded
    with 0 and cut by their length."""
    from dompap_amd.engine import Engine
    from dompap_amd.potential import device_potential
    rng = np.random.default_rng(17)
    L = 200.0
    dense = rng.uniform(10.0, 23.0, size=(13 * 13 * 400, 2))
    dilute = rng.uniform(0.0, L, size=(3000, 2))
    x = np.ascontiguousarray(np.concatenate([dense, dilute]))
    n = x.shape[0]
    box = np.array([L, L])
    K = 1600
    eng = Engine(2, n, 1, "mixed")
    eng.set_box(box)
    dp = device_potential("(1-r)**2", 1.0)
    eng.load_potential(dp.code, 1.0, dp.u_rc, [[1.0]], [[1.0]])
    eng.set_neighbor_params(1.0, 0.0, K, 1)
    eng.set_field("positions", x)
    plan0 = eng.plan()
    assert plan0["entry_shift"] == 3 and plan0["sentinel_padding"] == 0
    eng.build_neighbor_list()
    assert eng.plan()["entry_bits"] == 32
    nlo, cnt = C.neighbor_list(x, box, 1.0, K, "cells")
    assert np.array_equal(eng.neighbor_counts(), cnt)
    F = eng.compute(forces=True)["forces"]
    Fo = C.pair(x, box, nlo, "harm2", 1.0)["forces"]
    frms = float(np.sqrt(np.mean(np.sum(Fo ** 2, axis=1))))
    assert np.abs(F - Fo).max() / frms < 1e-5
    nl = eng.neighbor_list()
    for i in rng.choice(dense.shape[0], 200, replace=False):
        assert set(nl[i][nl[i] >= 0].tolist()) == set(nlo[i][nlo[i] >= 0].tolist())
    eng.close()


@pytest.mark.parametrize("precision", ["mixed", "fp64"])
def test_entry_format_and_sentinel_padding(precision):
    """3-D Lennard-Jones liquid-like state (perturbed FCC, N = 4,000, rho =
    0.8442): mixed precision addresses 16-byte staged records by byte offset
    (shift 4) and pads rows with the sentinel's entry; FP64 mode keeps staged
    indices padded with 0.  Lists and forces match the C oracle either way,
    also after steps that leave rows of every length modulo the chunk."""
    from dompap_amd.engine import Engine
    from dompap_amd.potential import device_potential
    rng = np.random.default_rng(5)
    a = (4.0 / 0.8442) ** (1.0 / 3.0)
    cells = np.stack(np.meshgrid(*[np.arange(10)] * 3, indexing="ij"), -1).reshape(-1, 3)
    basis = np.array([[0, 0, 0], [0.5, 0.5, 0], [0.5, 0, 0.5], [0, 0.5, 0.5]])
    x = ((cells[:, None, :] + basis[None]) * a).reshape(-1, 3)
    L = 10 * a
    x = np.mod(x + rng.normal(0.0, 0.15, x.shape), L)
    n = x.shape[0]
    box = np.array([L, L, L])
    eng = Engine(3, n, 1, precision)
    eng.set_box(box)
    dp = device_potential("4*(r**-12 - r**-6)", 2.5)
    eng.load_potential(dp.code, 2.5, dp.u_rc, [[1.0]], [[1.0]])
    eng.set_neighbor_params(3.0, 0.5, 256, 1)
    eng.set_field("positions", x)
    eng.build_neighbor_list()
    plan = eng.plan()
    if precision == "mixed":
        assert plan["entry_shift"] == 4 and plan["sentinel_padding"] == 1
    else:
        assert plan["entry_shift"] == 0 and plan["sentinel_padding"] == 0
    nlo, cnt = C.neighbor_list(x, box, 3.0, 256, "cells")
    assert np.array_equal(eng.neighbor_counts(), cnt)
    assert len(set((cnt % 4).tolist())) == 4  # rows end at every position of a chunk
    F = eng.compute(forces=True)["forces"]
    Fo = C.pair(x, box, nlo, "lj", 2.5)["forces"]
    frms = float(np.sqrt(np.mean(np.sum(Fo ** 2, axis=1))))
    tol = 1e-5 if precision == "mixed" else 1e-10
    assert np.abs(F - Fo).max() / frms < tol
    eng.close()
