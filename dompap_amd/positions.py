"""Host-side position helpers with the reference's API (dompap/positions.py).

These are set-up and analysis utilities (lattice generation, wrapping a host
array, the radial distribution function).  The per-step path -- minimum
image, wrap after each integrator step, neighbour search -- runs on the GPU
inside libdompap_hip.so.
"""
from __future__ import annotations

import math

import numpy as np


def get_displacement_vector(position, other_position, box_vectors) -> np.ndarray:
    """positions.py:6-17: d = a - b, one +-L shift per dimension."""
    d = np.array(position, dtype=np.float64) - np.array(other_position, dtype=np.float64)
    for k in range(d.shape[-1]):
        if d[k] > box_vectors[k] / 2:
            d[k] -= box_vectors[k]
        elif d[k] < -box_vectors[k] / 2:
            d[k] += box_vectors[k]
    return d


def get_distance(position, other_position, box_vectors) -> float:
    """positions.py:20-30."""
    d = get_displacement_vector(position, other_position, box_vectors)
    s = 0.0
    for v in d:
        s = s + v * v
    return np.sqrt(s)


def generate_positions(unit_cell_coordinates, cells, lattice_constants) -> np.ndarray:
    """positions.py:33-52: lattice sites, cells in C order, then scaled."""
    ucc = np.asarray(unit_cell_coordinates, dtype=np.float64)
    cells = np.asarray(cells, dtype=np.int64)
    lc = np.asarray(lattice_constants, dtype=np.float64)
    ncell = int(np.prod(cells))
    idx = np.array(np.unravel_index(np.arange(ncell), tuple(int(c) for c in cells)), dtype=np.float64).T
    pos = (ucc[None, :, :] + idx[:, None, :]).reshape(-1, ucc.shape[1])
    return pos * lc


def wrap_into_box(positions: np.ndarray, image_positions: np.ndarray, box_vectors: np.ndarray) -> None:
    """positions.py:90-98 (in place)."""
    shift = np.floor(positions / box_vectors[np.newaxis, :])
    image_positions += shift.astype(np.int32)
    positions -= shift * box_vectors[np.newaxis, :]


def get_radial_distribution_function(positions_n, positions_m, box_vectors, r_bins):
    """positions.py:58-87: histogram of minimum-image distances normalised by
    the n-sphere shell volume."""
    r_bins = np.asarray(r_bins, dtype=np.float64)
    if not np.all(np.diff(r_bins) >= 0):
        raise ValueError("r_bins array must be sorted in ascending order.")
    diff = positions_n[:, np.newaxis, :] - positions_m[np.newaxis, :, :]
    for i in range(diff.shape[-1]):
        box_length = np.linalg.norm(box_vectors[i])
        diff[:, :, i] -= box_length * np.round(diff[:, :, i] / box_length)
    distances = np.linalg.norm(diff, axis=2)
    histogram = np.histogram(distances, bins=r_bins)[0]
    return normalise_radial_histogram(histogram, r_bins, positions_n.shape[0] * positions_m.shape[0],
                                      box_vectors, positions_n.shape[1])


def normalise_radial_histogram(histogram, r_bins, number_of_pairs, box_vectors, dimensions):
    """positions.py:79-87: bin centres and g(r) = counts / shell volume /
    number of pairs * box volume, the shell volume being the n-sphere
    surface 2 pi^(D/2) / Gamma(D/2) r_mid^(D-1) times the bin width."""
    box_volume = np.prod(box_vectors)
    dr_bins = np.diff(r_bins)
    r_mid = r_bins[:-1] + dr_bins / 2
    surface = 2 * np.pi ** (dimensions / 2) / math.gamma(dimensions / 2)
    shell_volume = surface * r_mid ** (dimensions - 1) * dr_bins
    return r_mid, histogram / shell_volume / number_of_pairs * box_volume
