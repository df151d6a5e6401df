"""Batched replicas: many independent systems stepped together on one GPU.

The reference runs one ``Simulation`` per process.  ``ReplicaBatch`` takes a
list of ``Simulation`` objects of one dimension (and up to 8 distinct pair
potentials / parameter tables, each replica computing with its own) and runs
them as the replicas of one device context: one launch per kernel for
all of them, 100-step hipGraph blocks, per-replica boxes, cutoffs, skins,
time steps, temperatures and damping times.  Each replica's noise is keyed by
its own seed and the particle index inside the replica, so replica r draws
the numbers a single ``Simulation`` with that seed would draw.

Rebuilds are per replica: before every step the block decides, for each
replica, whether its own skin trigger fired (neighbor_list.py:183-201) and
rebuilds only those replicas' lists; each replica's update counter counts its
own rebuilds, as a single run would.
"""
from __future__ import annotations

from typing import Sequence

import numpy as np

from . import _lib
from .engine import Engine
from .potential import device_potential, lower_parameters


class ReplicaBatch:
    def __init__(self, simulations: Sequence, precision: str | None = None, device: int = 0,
                 seeds: Sequence[int] | None = None, pair_precision: str | None = None, block_steps: int = 100):
        sims = list(simulations)
        if not sims:
            raise ValueError("ReplicaBatch needs at least one Simulation")
        self.simulations = sims
        s0 = sims[0]
        dim = s0.get_dimensions_of_space()
        tabs = []
        for s in sims:
            if s.get_dimensions_of_space() != dim:
                raise ValueError("all replicas must have the same dimension of space")
            if s.pair_potential_str is None:
                raise ValueError("every replica needs a pair potential (set_pair_potential)")
            types = np.asarray(s.particle_types).reshape(-1)
            if types.shape[0] != s.number_of_particles():
                types = np.zeros(s.number_of_particles(), dtype=np.int32)
            tabs.append(lower_parameters(s.sigma_func, s.epsilon_func, types,
                                         s.__dict__.get("_sigma_const"), s.__dict__.get("_epsilon_const")))
        # one device potential per distinct (expression, r_cut, sigma / epsilon
        # tables); the type tables share one size (padded with unused types)
        nt = max(t.sigma.shape[0] for t in tabs)

        def pad(a):
            out = np.ones((nt, nt))
            out[: a.shape[0], : a.shape[1]] = a
            return out

        keys, rep_pot = [], []
        for s, t in zip(sims, tabs):
            key = (s.pair_potential_str, float(s.pair_potential_r_cut), pad(t.sigma).tobytes(),
                   pad(t.epsilon).tobytes())
            if key not in keys:
                keys.append(key)
            rep_pot.append(keys.index(key))
        if len(keys) > 8:
            raise ValueError(f"a batch holds at most 8 distinct potentials / parameter tables, got {len(keys)}")
        self.rep_pot = np.asarray(rep_pot, dtype=np.int32)
        self.dim = dim
        self.counts = np.array([s.number_of_particles() for s in sims], dtype=np.int64)
        self.offsets = np.concatenate([[0], np.cumsum(self.counts)]).astype(np.int64)
        n = int(self.offsets[-1])
        R = len(sims)
        prec = precision if precision is not None else s0.__dict__.get("_precision", "fp64")
        pp = pair_precision if pair_precision is not None else s0.__dict__.get("_pair_precision", "fp64")
        eng = Engine(dim, n, R, prec, device)
        self.engine = eng
        boxes = np.stack([np.asarray(s.box_vectors, dtype=np.float64).reshape(dim) for s in sims])
        eng.set_replicas(self.offsets, boxes)
        dps = [device_potential(k[0], k[1]) for k in keys]
        first = [rep_pot.index(p) for p in range(len(keys))]
        eng.load_potentials([d.code for d in dps], [d.r_cut for d in dps], [d.u_rc for d in dps],
                            [pad(tabs[i].sigma) for i in first], [pad(tabs[i].epsilon) for i in first],
                            self.rep_pot, pp)
        # Simulation.py:353-356 per replica
        cut = np.array([t.sigma_max * float(s.pair_potential_r_cut) + float(s.neighbor_list_skin)
                        for s, t in zip(sims, tabs)])
        skin = np.array([float(s.neighbor_list_skin) for s in sims])
        K = int(max(int(s.max_number_of_neighbors) for s in sims))
        order = _lib.DP_ORDER_X_FASTEST if dim == 3 else _lib.DP_ORDER_LAST_FASTEST
        eng.set_neighbor_params(cut, skin, K, order)
        eng.set_integrator([float(s.time_step) for s in sims], [float(s.temperature_target) for s in sims],
                           [float(s.temperature_damping_time) for s in sims])
        seeds = [int(s.seed) for s in sims] if seeds is None else [int(x) for x in seeds]
        if len(seeds) != R:
            raise ValueError("one seed per replica")
        eng.set_seed(seeds[0], int(s0.number_of_steps))
        eng.set_replica_seeds(seeds)
        eng.set_graph_block(int(block_steps))
        self.seeds = seeds

        def cat(name, width, dtype):
            parts = []
            for s in sims:
                a = np.asarray(getattr(s, name), dtype=dtype)
                want = (s.number_of_particles(), width) if width else (s.number_of_particles(),)
                if a.size != int(np.prod(want)):
                    raise ValueError(f"{name} of a replica has shape {a.shape}, expected {want}")
                parts.append(a.reshape(want))
            return np.ascontiguousarray(np.concatenate(parts))

        eng.set_field("positions", cat("positions", dim, np.float64))
        eng.set_field("velocities", cat("velocities", dim, np.float64))
        eng.set_field("betas", cat("betas", dim, np.float64))
        eng.set_field("image_positions", cat("image_positions", dim, np.int32))
        eng.set_field("masses", cat("masses", 0, np.float64))
        eng.set_field("types", np.ascontiguousarray(np.concatenate([t.types for t in tabs])))
        eng.build_neighbor_list()
        self.number_of_steps = int(s0.number_of_steps)
        # the build above is this context's own start, not a reference rebuild
        self._updates0 = np.array([s.number_of_neighbor_list_updates for s in sims], dtype=np.int64)

    # -- stepping -----------------------------------------------------------
    @property
    def n_replicas(self) -> int:
        return len(self.simulations)

    def run(self, steps: int = 1000) -> None:
        """``steps`` G-JF steps of every replica (Simulation.py:726-737)."""
        if steps > 0:
            self.engine.step(int(steps), "gjf")
            self.number_of_steps += int(steps)

    def step(self, noise=None) -> None:
        """One G-JF step of every replica; ``noise`` [N_total, D] optional."""
        self.engine.step(1, "gjf", xi=noise)
        self.number_of_steps += 1

    def step_leap_frog(self) -> None:
        self.engine.step(1, "leapfrog")
        self.number_of_steps += 1

    # -- observables ----------------------------------------------------------
    def _split(self, a):
        return [a[self.offsets[r]:self.offsets[r + 1]] for r in range(self.n_replicas)]

    def get_potential_energy(self) -> np.ndarray:
        """Per-replica potential energies (potential.py:69-100)."""
        return np.array(self.engine.compute(forces=False, energy=True)["energy"])

    def get_virial(self) -> np.ndarray:
        """Per-replica virials (potential.py:253-279 over the list)."""
        return np.array(self.engine.compute(forces=False, virial=True)["virial"])

    def get_forces(self) -> list:
        """Per-replica force arrays (potential.py:163-188)."""
        return self._split(self.engine.compute(forces=True)["forces"])

    def get_field(self, name: str) -> list:
        """Per-replica copies of a state array ("positions", "velocities", ...)."""
        return self._split(self.engine.get_field(name))

    def number_of_neighbor_list_updates(self) -> np.ndarray:
        upd, _, _ = self.engine.stats()
        return self._updates0 + upd

    def write_back(self) -> None:
        """Copy every replica's state back into its Simulation object."""
        fields = {f: self.get_field(f) for f in ("positions", "velocities", "betas", "image_positions")}
        upd = self.number_of_neighbor_list_updates()
        for r, s in enumerate(self.simulations):
            for f, parts in fields.items():
                setattr(s, f, np.array(parts[r]))
            s.__dict__["number_of_steps"] = self.number_of_steps
            s.number_of_neighbor_list_updates = int(upd[r])

    def close(self) -> None:
        self.engine.close()
