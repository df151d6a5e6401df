"""ReplicaBatch argument checks (host only: they run before any device call)."""
import pytest


def test_batch_rejects_inconsistent_replicas():
    from dompap_amd import ReplicaBatch, Simulation
    a = Simulation()
    b = Simulation()
    b.set_positions(([0.0, 0.0],), (5, 5), (1.0, 1.0))
    with pytest.raises(ValueError, match="dimension"):
        ReplicaBatch([a, b])
    # replicas may differ in potential, but a batch holds at most 8 distinct ones
    many = []
    for k in range(9):
        s = Simulation()
        s.set_pair_potential("4*(r**-12-r**-6)", r_cut=2.0 + 0.1 * k)
        many.append(s)
    with pytest.raises(ValueError, match="at most 8 distinct potentials"):
        ReplicaBatch(many)
    with pytest.raises(ValueError, match="at least one"):
        ReplicaBatch([])
