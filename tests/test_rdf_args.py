"""Argument checks of Simulation.get_radial_distribution_function that fire
before any device work (positions.py:58-87 raises on unsorted edges)."""
import numpy as np
import pytest


def test_rdf_rejects_unsorted_and_short_edges():
    from dompap_amd import Simulation
    s = Simulation()
    with pytest.raises(ValueError, match="ascending"):
        s.get_radial_distribution_function(np.array([0.0, 2.0, 1.0]))
    with pytest.raises(ValueError, match="ascending"):
        s.get_radial_distribution_function(np.array([0.0, np.nan, 1.0]))
    with pytest.raises(ValueError):
        s.get_radial_distribution_function(np.array([1.0]))
