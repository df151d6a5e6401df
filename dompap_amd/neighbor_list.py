"""Neighbour-list free functions with the reference's API
(dompap/neighbor_list.py), computed by the GPU cell-list pipeline.

Each call builds a short-lived one-replica context; ``Simulation`` keeps its
list on the device instead.
"""
from __future__ import annotations

import numpy as np

from . import _lib
from .engine import Engine


def _build(positions, box_vectors, cutoff_distance, max_number_of_neighbors, order):
    x = np.ascontiguousarray(np.asarray(positions, dtype=np.float64))
    n, dim = x.shape
    eng = Engine(dim, n, 1, "fp64")
    try:
        eng.set_box(np.asarray(box_vectors, dtype=np.float64))
        eng.set_neighbor_params(float(cutoff_distance), 0.0, int(max_number_of_neighbors), order)
        eng.set_field("positions", x)
        eng.build_neighbor_list()
        return eng.neighbor_list()
    finally:
        eng.close()


def _ascending(nl):
    big = np.iinfo(np.int32).max
    key = np.where(nl < 0, big, nl)
    key.sort(axis=1)
    return np.where(key == big, -1, key).astype(np.int32)


def get_neighbor_list_double_loop(positions, box_vectors, cutoff_distance, max_number_of_neighbors) -> np.ndarray:
    """neighbor_list.py:7-39: rows of ascending m with sqrt(sum d^2) < cutoff,
    -1 padded; ValueError if a row exceeds max_number_of_neighbors."""
    return _ascending(_build(positions, box_vectors, cutoff_distance, max_number_of_neighbors,
                             _lib.DP_ORDER_X_FASTEST))


def get_neighbor_list_cell_list(positions, box_vectors, cutoff_distance, max_number_of_neighbors) -> np.ndarray:
    """neighbor_list.py:42-103 (any dimension): rows in stencil order, last
    dimension fastest.  Same neighbour sets as the double loop (the
    reference's pairs lost to a narrow last cell are found here)."""
    return _build(positions, box_vectors, cutoff_distance, max_number_of_neighbors, _lib.DP_ORDER_LAST_FASTEST)


def get_neighbor_list_cell_list_3d(positions, box_vectors, cutoff_distance, max_number_of_neighbors) -> np.ndarray:
    """neighbor_list.py:106-180: 3-D only, stencil x fastest."""
    if np.asarray(positions).shape[1] != 3:
        raise NotImplementedError("get_neighbor_list_cell_list_3d is only for 3 spatial dimensions.")
    return _build(positions, box_vectors, cutoff_distance, max_number_of_neighbors, _lib.DP_ORDER_X_FASTEST)


def neighbor_list_is_old(positions, old_positions, box_vectors, skin) -> bool:
    """neighbor_list.py:183-201: max |minimage(x - x_ref)|^2 > (skin/2)^2,
    evaluated by the device skin check."""
    x = np.ascontiguousarray(np.asarray(positions, dtype=np.float64))
    old = np.ascontiguousarray(np.asarray(old_positions, dtype=np.float64))
    n, dim = x.shape
    eng = Engine(dim, n, 1, "fp64")
    try:
        eng.set_box(np.asarray(box_vectors, dtype=np.float64))
        eng.set_neighbor_params(0.0, float(skin), 1, _lib.DP_ORDER_X_FASTEST)
        eng.set_field("positions", old)
        eng.build_neighbor_list()
        eng.set_field("positions", x)
        return eng.update_neighbor_list()
    finally:
        eng.close()


def get_number_of_neighbors(neighbor_list) -> np.ndarray:
    """neighbor_list.py:204-207."""
    return np.sum(np.asarray(neighbor_list) != -1, axis=1)
