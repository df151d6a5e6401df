"""Batched replicas (ReplicaBatch) against single-replica runs of the same
systems: forces, energies and virials on equal positions, and short G-JF
trajectories with device noise (replica r keyed like a single run with its
seed).  FP64 mode: <= 1e-12 relative for static quantities (summation order
may differ with the tile plan), <= 1e-8 for positions after 25 steps (the
batch's tiles order the sums differently).  Rebuilds are per replica: each
replica rebuilds exactly when its single run does.  The two step-block
modes (one rebuild pipeline per block, or a gated one before every step)
give bit-identical trajectories."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu
FCC = [(0, 0, 0), (.5, .5, 0), (.5, 0, .5), (0, .5, .5)]


def _replica(precision, cells, density, T, seed, noise=0.05):
    from dompap_amd import Simulation
    s = Simulation(seed=seed, precision=precision)
    s.set_positions(unit_cell_coordinates=FCC, cells=(cells,) * 3)
    s.set_density(density)
    s.set_masses(1.0)
    rng = np.random.default_rng(seed)
    s.positions = np.asarray(s.positions) + rng.normal(0, noise, size=s.positions.shape)
    s.velocities = rng.normal(0, np.sqrt(T), size=s.positions.shape)
    s.betas = np.zeros(s.positions.shape)
    s.set_pair_potential("4*(r**-12-r**-6)", r_cut=2.5)
    s.set_neighbor_list(skin=0.5, max_number_of_neighbors=160)
    s.set_integrator(time_step=0.005, target_temperature=T, temperature_damping_time=1.0)
    return s


SPECS = [(5, 0.80, 0.9), (6, 0.90, 1.2), (7, 0.75, 0.6), (4, 1.00, 1.5)]


def _frms(F):
    return float(np.sqrt(np.mean(np.sum(F ** 2, axis=1))))


def test_batch_static_observables_match_single_replicas():
    from dompap_amd import ReplicaBatch
    sims = [_replica("fp64", c, rho, T, 100 + i) for i, (c, rho, T) in enumerate(SPECS)]
    singles = [s.copy() for s in sims]
    b = ReplicaBatch(sims)
    Fb, Eb, Wb = b.get_forces(), b.get_potential_energy(), b.get_virial()
    for r, s in enumerate(singles):
        F = s.get_forces()
        assert np.abs(Fb[r] - F).max() / _frms(F) < 1e-12
        assert abs(Eb[r] - s.get_potential_energy()) / abs(s.get_potential_energy()) < 1e-12
        assert abs(Wb[r] - s.get_virial()) / abs(s.get_virial()) < 1e-12
    b.close()


def test_batch_trajectories_match_single_replicas():
    from dompap_amd import ReplicaBatch
    sims = [_replica("fp64", c, rho, T, 200 + i) for i, (c, rho, T) in enumerate(SPECS)]
    singles = [s.copy() for s in sims]
    b = ReplicaBatch(sims)
    b.run(25)
    xb = b.get_field("positions")
    for r, s in enumerate(singles):
        s.run(25)
        x = np.asarray(s.positions)
        assert np.abs(xb[r] - x).max() < 1e-8, r
    b.write_back()
    for r, s in enumerate(sims):
        assert np.array_equal(np.asarray(s.positions), xb[r])
        assert s.number_of_steps == 25
    b.close()


def test_batch_mixed_many_uneven_replicas():
    """Config-3-like batch (64 replicas, N log-uniform in [500, 5000], density
    U[0.7, 1.0], T U[0.5, 2.0], perturbed lattices) in mixed precision, 200
    steps in graph blocks; forces of sampled replicas against FP64 single
    runs at the batch's positions."""
    from dompap_amd import ReplicaBatch, Simulation
    rng = np.random.default_rng(2024)
    sims = []
    for r in range(64):
        n_r = int(round(np.exp(rng.uniform(np.log(500), np.log(5000)))))
        c = int(np.ceil((n_r / 4) ** (1 / 3)))
        s = _replica("mixed", c, rng.uniform(0.7, 1.0), rng.uniform(0.5, 2.0), 1000 + r)
        keep = np.sort(rng.choice(s.number_of_particles(), size=min(n_r, s.number_of_particles()), replace=False))
        for f in ("positions", "velocities", "betas", "image_positions"):
            setattr(s, f, np.asarray(getattr(s, f))[keep])
        s.masses = np.ones((keep.size, 1))
        s.particle_types = np.zeros((keep.size, 1), dtype=np.int32)
        sims.append(s)
    b = ReplicaBatch(sims, precision="mixed")
    b.run(200)
    xb = b.get_field("positions")
    Fb = b.get_forces()
    E = b.get_potential_energy()
    assert np.all(np.isfinite(E))
    for r in (0, 17, 41, 63):
        ref = Simulation(seed=1, precision="fp64")
        ref.box_vectors = np.asarray(sims[r].box_vectors)
        ref.positions = xb[r]
        ref.velocities = np.zeros_like(xb[r])
        ref.betas = np.zeros_like(xb[r])
        ref.image_positions = np.zeros(xb[r].shape, dtype=np.int32)
        ref.masses = np.ones((xb[r].shape[0], 1))
        ref.particle_types = np.zeros((xb[r].shape[0], 1), dtype=np.int32)
        ref.set_pair_potential("4*(r**-12-r**-6)", r_cut=2.5)
        ref.set_neighbor_list(skin=0.5, max_number_of_neighbors=160)
        F = ref.get_forces()
        assert np.abs(Fb[r] - F).max() / _frms(F) < 1e-5, r
    b.close()


def test_batch_rebuilds_per_replica_like_single_runs():
    from dompap_amd import ReplicaBatch
    specs = [(5, 0.80, 0.4), (6, 0.90, 2.0), (5, 0.75, 0.8), (4, 1.00, 1.5)]
    sims = [_replica("fp64", c, rho, T, 300 + i) for i, (c, rho, T) in enumerate(specs)]
    singles = [s.copy() for s in sims]
    b = ReplicaBatch(sims, block_steps=100)
    b.run(150)
    xb = b.get_field("positions")
    upd = b.number_of_neighbor_list_updates()
    counts = []
    for r, s in enumerate(singles):
        s.run(150)
        counts.append(s.number_of_neighbor_list_updates)
        assert np.abs(xb[r] - np.asarray(s.positions)).max() < 1e-8, r
    assert list(upd) == counts
    # hot and cold replicas rebuild at different rates
    assert len(set(counts)) > 1
    b.close()


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_block_rebuild_modes_give_identical_trajectories(precision):
    runs = []
    for mode, block in ((0, 16), (1, 16), (1, 100), (0, 0)):
        s = _replica(precision, 6, 0.85, 1.2, 77)
        s.run(1)  # creates the context
        s._engine.set_block_rebuild(mode)
        s._engine.set_graph_block(block)
        s.run(130)
        runs.append((np.asarray(s.positions).copy(), np.asarray(s.velocities).copy(),
                     s.number_of_neighbor_list_updates))
    for x, v, u in runs[1:]:
        assert np.array_equal(x, runs[0][0]) and np.array_equal(v, runs[0][1]) and u == runs[0][2]


POTS = {"lj": ("4*(r**-12-r**-6)", 2.5), ("ipl", 12): ("r**-12", 1.5), "harm2": ("(1-r)**2", 1.0)}


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_batch_of_different_potentials_against_c_oracle(precision):
    """Replicas of three potentials (LJ, r^-12, (1-r)^2) stepped in the same
    launches, each computing with its own potential: forces / energy / virial
    of every replica against the C oracle with that potential."""
    from dompap_amd import ReplicaBatch
    from oracle import c_oracle as C
    plan = [("lj", 5, 0.85, 1.0), ("harm2", 6, 1.2, 0.01), (("ipl", 12), 5, 1.0, 1.0), ("lj", 4, 0.95, 0.7),
            (("ipl", 12), 4, 0.9, 1.5), ("harm2", 5, 1.1, 0.02), ("lj", 6, 0.8, 1.2)]
    sims = []
    for i, (pot, cells, rho, T) in enumerate(plan):
        s = _replica(precision, cells, rho, T, 500 + i)
        expr, rc = POTS[pot]
        s.set_pair_potential(expr, r_cut=rc)
        s.set_neighbor_list(skin=0.4, max_number_of_neighbors=160)
        sims.append(s)
    b = ReplicaBatch(sims)
    assert sorted(set(b.rep_pot.tolist())) == [0, 1, 2]
    b.run(30)
    xs = b.get_field("positions")
    Fb, Eb, Wb = b.get_forces(), b.get_potential_energy(), b.get_virial()
    tol_f, tol_e = (1e-10, 1e-10) if precision == "fp64" else (1e-5, 1e-6)
    for r, (pot, cells, rho, T) in enumerate(plan):
        x = xs[r]
        box = np.asarray(sims[r].box_vectors, dtype=np.float64)
        rc = POTS[pot][1]
        nl, _ = C.neighbor_list(x, box, rc + 0.4, 160, "cells")
        o = C.pair(x, box, nl, pot, rc, energy=True, virial=True)
        n = x.shape[0]
        assert np.abs(Fb[r] - o["forces"]).max() / _frms(o["forces"]) < tol_f, r
        assert abs(Eb[r] - o["energy"]) < tol_e * max(abs(o["energy"]), n), r
        assert abs(Wb[r] - o["virial"]) < 10 * tol_e * max(abs(o["virial"]), n), r
    b.close()
