"""Simulation: drop-in replacement for dompap's Simulation class
(dompap/Simulation.py:67-1229) whose per-step path runs on an MI355X.

Same constructor defaults, method names, argument meaning, attributes,
method strings and error behaviour as the reference.  What changes is where
the state lives: positions, velocities, betas, image positions and the
neighbour list are held on the GPU (cell-sorted) by a libdompap_hip context,
and ``step()`` / ``run()`` / ``get_forces()`` / ``get_potential_energy()`` /
``get_virial()`` / ``get_laplacian()`` launch HIP kernels.  Reading an
attribute such as ``sim.positions`` downloads a host copy; writing into that
copy (``sim.positions[0] = ...``) or assigning the attribute is uploaded
before the next device call.

Additions (keyword arguments, defaults keep the reference's behaviour):
``precision`` ("fp64" or "mixed"), ``device``, ``seed`` (Langevin noise
stream: Philox4x32-10 keyed by seed, step and particle), and ``noise=`` on
``step()`` to inject the standard normals of integrator.py:30.
"""
from __future__ import annotations

import copy as _copy
import math
import os
import weakref

import numpy as np

from . import _lib
from .engine import Engine
from .potential import (device_potential, hardcoded_pair_potentials, lower_parameters,
                        make_pair_potential)

DEFAULT_SPATIAL_DIMENSION = 3
DEFAULT_NUMBER_OF_PARTICLES = 125  # 5x5x5 simple cubic lattice


def default_particle_types():
    return np.zeros(shape=(DEFAULT_NUMBER_OF_PARTICLES, 1), dtype=np.int32)


def default_positions():
    """Simulation.py:17-27: 5x5x5 simple cubic lattice, last index fastest."""
    idx = np.arange(DEFAULT_NUMBER_OF_PARTICLES)
    return np.stack([idx // 25, (idx // 5) % 5, idx % 5], axis=1).astype(np.float64)


def default_box_vectors():
    return np.array([5, 5, 5], dtype=np.float64)


def default_image_positions():
    return np.zeros(shape=(DEFAULT_NUMBER_OF_PARTICLES, DEFAULT_SPATIAL_DIMENSION), dtype=np.int32)


def default_velocities():
    return np.random.normal(loc=0.0, scale=1.0, size=(DEFAULT_NUMBER_OF_PARTICLES, DEFAULT_SPATIAL_DIMENSION))


def default_betas():
    return np.zeros(shape=(DEFAULT_NUMBER_OF_PARTICLES, DEFAULT_SPATIAL_DIMENSION), dtype=np.float64)


def default_masses():
    return np.ones(shape=(DEFAULT_NUMBER_OF_PARTICLES, 1), dtype=np.float64)


class _Mirror(np.ndarray):
    """Host copy of a device-backed field; in-place writes mark the field for
    upload (only while this copy is the simulation's current copy)."""

    def __array_finalize__(self, obj):
        self._dp_owner = getattr(obj, "_dp_owner", None)

    def __array_wrap__(self, arr, context=None, return_scalar=False):
        out = np.asarray(arr).view(np.ndarray)
        if return_scalar and out.ndim == 0:
            return out[()]
        return out

    def _touch(self):
        own = self._dp_owner
        if own is not None:
            sim = own[0]()
            if sim is not None:
                sim._host_written(own[1], own[2])

    def __setitem__(self, key, value):
        super().__setitem__(key, value)
        self._touch()

    def copy(self, order="C"):
        return np.array(self, order=order, copy=True).view(np.ndarray)

    def __deepcopy__(self, memo):
        return self.copy()

    def __reduce__(self):
        return np.asarray(self).view(np.ndarray).__reduce__()


def _inplace(name):
    base = getattr(np.ndarray, name)

    def op(self, other):
        res = base(self, other)
        self._touch()
        return res

    return op


for _n in ("__iadd__", "__isub__", "__imul__", "__itruediv__", "__ifloordiv__", "__ipow__", "__imod__",
           "__iand__", "__ior__", "__ixor__", "__ilshift__", "__irshift__"):
    setattr(_Mirror, _n, _inplace(_n))

# fields whose authoritative copy may live on the device
_DEVICE_FIELDS = ("positions", "velocities", "betas", "image_positions")
# host-owned fields uploaded when changed
_HOST_FIELDS = ("particle_types", "masses", "box_vectors")


class Simulation:
    """The Simulation class (dompap/Simulation.py:67-147) on an MI355X.

    Examples
    --------

        >>> from dompap_amd import Simulation
        >>> sim = Simulation()
        >>> sim.set_positions(unit_cell_coordinates=([0.0, 0.0, 0.0],), cells=(5, 5, 5), lattice_constants=(1.0, 1.0, 1.0))
        >>> sim.set_masses(masses=1.0)
        >>> sim.set_random_velocities(temperature=1.0)
        >>> sim.set_pair_potential(pair_potential_str='(1-r)**2', r_cut=1.0)
        >>> sim.set_neighbor_list(skin=1.0, max_number_of_neighbors=128)
        >>> sim.set_integrator(time_step=0.01, target_temperature=1.0, temperature_damping_time=1.0)
    """

    # Selecting algorithms (Simulation.py:122-133)
    energy_method_str = "neighbor list"
    force_method_str = "neighbor list"
    neighbor_list_method_str = "double loop"
    _KNOWN_ENERGY_METHODS = {"neighbor list", "double loop", "double loop single core"}
    _KNOWN_FORCE_METHODS = {"neighbor list", "double loop", "double loop single core", "vectorized"}
    _KNOWN_CELL_LIST_METHODS = {"cell list", "double loop"}

    def __init__(self, particle_types=None, positions=None, box_vectors=None, image_positions=None,
                 velocities=None, betas=None, neighbor_list=None, positions_neighbour_list=None,
                 masses=None, pair_potential_r_cut=np.float64(1.0), neighbor_list_skin=np.float64(1.0),
                 max_number_of_neighbors=np.int32(512), number_of_neighbor_list_updates=0,
                 number_of_steps=0, time_step=np.float64(0.01), temperature_target=np.float64(1.0),
                 temperature_damping_time=np.float64(0.1), pair_potential=None, pair_potential_str=None,
                 pair_force=None, pair_curvature=None, epsilon_func=None, sigma_func=None,
                 precision="fp64", device=0, seed=None, pair_precision="fp64"):
        d = self.__dict__
        d["_engine"] = None
        d["_h"] = {}
        d["_stale"] = set()
        d["_dirty"] = set()
        d["_tok"] = {}
        d["_precision"] = precision
        d["_pair_precision"] = pair_precision
        d["_device_index"] = int(device)
        d["seed"] = int(np.random.randint(0, 2**62)) if seed is None else int(seed)
        d["_params_dirty"] = True
        d["_integ_dirty"] = True
        d["_pot_dirty"] = True
        d["_tuning"] = (0, 0, -1)  # kernel shape: planner's choice
        d["_tuning_dirty"] = True
        d["_tables"] = None
        d["_nl_pending"] = False
        d["_nl_host_count"] = 0
        d["_dev_count_base"] = 0
        d["_dev_steps_base"] = 0
        d["_restore_xref"] = None
        self.particle_types = default_particle_types() if particle_types is None else particle_types
        self.positions = default_positions() if positions is None else positions
        self.box_vectors = default_box_vectors() if box_vectors is None else box_vectors
        self.image_positions = default_image_positions() if image_positions is None else image_positions
        self.velocities = default_velocities() if velocities is None else velocities
        self.betas = default_betas() if betas is None else betas
        self.masses = default_masses() if masses is None else masses
        d["pair_potential_r_cut"] = np.float64(pair_potential_r_cut)
        d["neighbor_list_skin"] = np.float64(neighbor_list_skin)
        d["max_number_of_neighbors"] = np.int32(max_number_of_neighbors)
        d["number_of_steps"] = int(number_of_steps)
        d["time_step"] = np.float64(time_step)
        d["temperature_target"] = np.float64(temperature_target)
        d["temperature_damping_time"] = np.float64(temperature_damping_time)
        d["pair_potential"] = pair_potential
        d["pair_potential_str"] = pair_potential_str
        d["pair_force"] = pair_force
        d["pair_curvature"] = pair_curvature
        d["epsilon_func"] = epsilon_func if epsilon_func is not None else (lambda n, m: np.float64(1.0))
        d["sigma_func"] = sigma_func if sigma_func is not None else (lambda n, m: np.float64(1.0))
        d["_sigma_const"] = 1.0 if sigma_func is None else None
        d["_epsilon_const"] = 1.0 if epsilon_func is None else None
        self._post_init()
        self.__dict__["_nl_host_count"] = int(number_of_neighbor_list_updates) + self._nl_host_count

    def _post_init(self):
        """Simulation.py:149-153."""
        self.set_pair_potential()
        self.set_pair_potential_parameters(sigma=1.0, epsilon=1.0)
        self.set_neighbor_list()

    # ------------------------------------------------------------------
    # attribute plumbing
    # ------------------------------------------------------------------
    def __getattr__(self, name):
        d = self.__dict__
        if name in _DEVICE_FIELDS or name in _HOST_FIELDS:
            return self._get_field(name)
        if name == "neighbor_list":
            return self._get_neighbor_list()
        if name == "positions_neighbour_list":
            return self._get_xref()
        if name == "number_of_neighbor_list_updates":
            return self._nl_updates()
        raise AttributeError(name)

    def __setattr__(self, name, value):
        if name in _DEVICE_FIELDS or name in _HOST_FIELDS:
            self._set_field(name, value)
            return
        if name == "number_of_neighbor_list_updates":
            self.__dict__["_nl_host_count"] = int(value) - self._device_auto_updates()
            return
        if name == "neighbor_list":
            return  # rebuilt on the device; assignment has no effect on the GPU list
        if name == "positions_neighbour_list":
            self.__dict__["_restore_xref"] = np.array(value, dtype=np.float64)
            return
        if name in ("pair_potential_r_cut", "neighbor_list_skin", "max_number_of_neighbors",
                    "neighbor_list_method_str"):
            self.__dict__["_params_dirty"] = True
        if name in ("time_step", "temperature_target", "temperature_damping_time"):
            self.__dict__["_integ_dirty"] = True
        if name in ("sigma_func", "epsilon_func"):
            self.__dict__["_pot_dirty"] = True
            self.__dict__["_params_dirty"] = True
            self.__dict__["_sigma_const" if name == "sigma_func" else "_epsilon_const"] = None
        self.__dict__[name] = value

    def _mirror(self, name, arr):
        m = np.ascontiguousarray(arr).view(_Mirror)
        tok = object()
        self._tok[name] = tok
        m._dp_owner = (weakref.ref(self), name, tok)
        return m

    def _host_written(self, name, tok):
        if self._tok.get(name) is tok:
            self._dirty.add(name)
            if name in ("particle_types", "box_vectors"):
                self.__dict__["_params_dirty"] = True
                self.__dict__["_pot_dirty"] = True

    def _set_field(self, name, value):
        if name == "particle_types":
            arr = np.array(value, dtype=np.int32)
        elif name == "image_positions":
            arr = np.array(value)
        else:
            arr = np.array(value, dtype=np.float64) if not isinstance(value, np.ndarray) else np.array(value)
        self._h[name] = self._mirror(name, arr)
        self._stale.discard(name)
        self._dirty.add(name)
        if name in ("particle_types", "box_vectors"):
            self.__dict__["_params_dirty"] = True
            self.__dict__["_pot_dirty"] = True

    def _get_field(self, name):
        if name in self._stale:
            eng = self._engine
            if name in ("image_positions",):
                arr = eng.get_field(name)
            else:
                arr = eng.get_field(name)
            self._h[name] = self._mirror(name, arr)
            self._stale.discard(name)
        return self._h[name]

    # ------------------------------------------------------------------
    # device synchronisation
    # ------------------------------------------------------------------
    def _shape(self):
        p = self._get_field("positions")
        return p.shape[0], p.shape[1]

    def _device(self) -> Engine:
        """Create / refresh the engine and upload everything the host changed."""
        n, dim = self._shape()
        eng = self._engine
        if eng is None or eng.n != n or eng.dim != dim:
            if eng is not None:
                # keep device-owned state before dropping the context
                for f in list(self._stale):
                    self._get_field(f)
                if self._restore_xref is None and n == eng.n and dim == eng.dim:
                    self.__dict__["_restore_xref"] = eng.get_field("positions_neighbour_list")
                self.__dict__["_nl_host_count"] = self._nl_updates()
                eng.close()
            eng = Engine(dim, n, 1, self._precision, self._device_index)
            self.__dict__["_engine"] = eng
            self.__dict__["_tuning_dirty"] = True
            self.__dict__["_dev_count_base"] = 0
            eng.set_seed(self.seed, self.number_of_steps)
            self._dirty.update(_DEVICE_FIELDS + _HOST_FIELDS)
            self.__dict__["_params_dirty"] = True
            self.__dict__["_integ_dirty"] = True
            self.__dict__["_pot_dirty"] = True
            if self._restore_xref is None and not self._nl_pending:
                self.__dict__["_nl_pending"] = True
        if "box_vectors" in self._dirty:
            box = np.asarray(self._h["box_vectors"], dtype=np.float64).reshape(-1)
            if box.shape[0] != dim:
                raise ValueError(f"box_vectors has {box.shape[0]} entries for {dim}-dimensional positions")
            eng.set_box(box)
            self._dirty.discard("box_vectors")
            self.__dict__["_params_dirty"] = True
        if self._pot_dirty or "particle_types" in self._dirty:
            self._refresh_potential(eng)
        if self._params_dirty:
            self._refresh_params(eng)
        if self._integ_dirty:
            eng.set_integrator(float(self.time_step), float(self.temperature_target),
                               float(self.temperature_damping_time))
            self.__dict__["_integ_dirty"] = False
        if self.__dict__.get("_tuning_dirty", True):
            eng.set_tuning(*self.__dict__.get("_tuning", (0, 0, -1)))
            eng.set_cell_scale(self.__dict__.get("_cell_scale", 1.0))
            self.__dict__["_tuning_dirty"] = False
        bad = self.__dict__.setdefault("_bad_fields", {})
        for f in ("positions", "velocities", "betas", "image_positions", "masses"):
            if f in self._dirty:
                v = np.asarray(self._h[f])
                want = (n,) if f == "masses" else (n, dim)
                if v.size != int(np.prod(want)):
                    if f == "positions":
                        raise ValueError(f"positions has shape {v.shape}, expected {want}")
                    # only the integrator reads these; as in the reference, a
                    # stale array of another shape is an error only when stepping
                    bad[f] = v.shape
                    continue
                bad.pop(f, None)
                v = v.astype(np.int32 if f == "image_positions" else np.float64).reshape(want)
                eng.set_field(f, v)
                self._dirty.discard(f)
        if self._restore_xref is not None:
            xr = np.asarray(self._restore_xref, dtype=np.float64)
            self.__dict__["_restore_xref"] = None
            if xr.shape == (n, dim) and not self._nl_pending:
                # rebuild the list at the reference positions, then restore
                eng.set_field("positions", xr)
                eng.build_neighbor_list()
                eng.set_field("positions", np.asarray(self._h["positions"], dtype=np.float64))
            else:
                self.__dict__["_nl_pending"] = True
        if self._nl_pending:
            eng.build_neighbor_list()
            self.__dict__["_nl_pending"] = False
        return eng

    def _refresh_potential(self, eng):
        types = np.asarray(self._h["particle_types"]).reshape(-1)
        if types.shape[0] != eng.n:
            # the reference never reads particle_types in its kernels (sigma and
            # epsilon are functions of the indices), so a stale types array of
            # another length is not an error there; treat it as one type
            types = np.zeros(eng.n, dtype=np.int32)
        tabs = lower_parameters(self.sigma_func, self.epsilon_func, types, self._sigma_const, self._epsilon_const)
        self.__dict__["_tables"] = tabs
        dp = device_potential(self.pair_potential_str, float(self.pair_potential_r_cut))
        eng.load_potential(dp.code, dp.r_cut, dp.u_rc, tabs.sigma, tabs.epsilon, self._pair_precision)
        eng.set_field("types", tabs.types)
        self._dirty.discard("particle_types")
        self.__dict__["_pot_dirty"] = False
        self.__dict__["_params_dirty"] = True

    def _refresh_params(self, eng):
        tabs = self._tables
        smax = tabs.sigma_max if tabs is not None else 1.0
        cutoff = smax * float(self.pair_potential_r_cut) + float(self.neighbor_list_skin)  # Simulation.py:353-356
        order = _lib.DP_ORDER_X_FASTEST if eng.dim == 3 else _lib.DP_ORDER_LAST_FASTEST
        eng.set_neighbor_params(cutoff, float(self.neighbor_list_skin), int(self.max_number_of_neighbors), order)
        self.__dict__["_params_dirty"] = False
        # new cells / capacity: the list is rebuilt at the next device call
        self.__dict__["_nl_pending"] = True

    def _device_auto_updates(self) -> int:
        eng = self._engine
        if eng is None:
            return 0
        upd, _, _ = eng.stats()
        return int(upd[0]) - self._dev_count_base

    def _nl_updates(self) -> int:
        return self._nl_host_count + self._device_auto_updates()

    def _after_device_step(self):
        self._stale.update(_DEVICE_FIELDS)
        for f in _DEVICE_FIELDS:
            self._tok.pop(f, None)
            self._h.pop(f, None)

    def _get_neighbor_list(self):
        eng = self._device()
        nl = eng.neighbor_list()
        if self.neighbor_list_method_str == "double loop":
            # neighbor_list.py:21-35 order: ascending m
            big = np.iinfo(np.int32).max
            key = np.where(nl < 0, big, nl)
            key.sort(axis=1)
            nl = np.where(key == big, -1, key).astype(np.int32)
        return nl

    def _get_xref(self):
        if self._engine is None:
            return np.array(self._get_field("positions"))
        return self._device().get_field("positions_neighbour_list")

    # ------------------------------------------------------------------
    # reference API
    # ------------------------------------------------------------------
    def __str__(self):
        return f"""Simulation:
    positions: {self.positions[:1]} ... {self.positions[-1:]}
    box_vectors: {self.box_vectors}
    masses: {self.masses[:1]} ... {self.masses[-1:]}"""

    def __deepcopy__(self, memo):
        for f in list(self._stale):
            self._get_field(f)
        new = object.__new__(type(self))
        nd = new.__dict__
        for k, v in self.__dict__.items():
            if k in ("_engine", "_h", "_stale", "_dirty", "_tok"):
                continue
            nd[k] = _copy.deepcopy(v, memo) if not callable(v) else v
        nd["_engine"] = None
        nd["_h"] = {}
        nd["_stale"] = set()
        nd["_dirty"] = set()
        nd["_tok"] = {}
        for k, v in self._h.items():
            new._set_field(k, np.array(v))
        nd["_nl_host_count"] = self._nl_updates()
        nd["_params_dirty"] = True
        nd["_integ_dirty"] = True
        nd["_pot_dirty"] = True
        nd["_tuning_dirty"] = True
        if self._engine is not None and not self._nl_pending:
            nd["_restore_xref"] = self._engine.get_field("positions_neighbour_list")
            nd["_nl_pending"] = False
        return new

    def copy(self):
        """Simulation.py:162-187: deep copy (device state included)."""
        return _copy.deepcopy(self)

    def set_positions(self, unit_cell_coordinates: tuple = ([0.0, 0.0, 0.0],), cells: tuple = (5, 5, 5),
                      lattice_constants: tuple = (1.0, 1.0, 1.0)):
        """Simulation.py:189-230."""
        from .positions import generate_positions

        unit_cell_coordinates = np.array(unit_cell_coordinates, dtype=np.float64)
        cells = np.array(cells, dtype=np.int32)
        lattice_constants = np.array(lattice_constants, dtype=np.float64)
        self.box_vectors = np.array(cells, dtype=np.float64) * np.array(lattice_constants, dtype=np.float64)
        self.positions = generate_positions(unit_cell_coordinates, cells, lattice_constants)
        pos = self._h["positions"]
        self.image_positions = np.zeros(shape=(pos.shape[0], pos.shape[1]), dtype=np.int32)
        if self.betas.shape != pos.shape:
            self.betas = np.zeros(shape=pos.shape, dtype=np.float64)
        if self.velocities.shape != pos.shape:
            self.velocities = np.random.normal(loc=0.0, scale=1.0, size=pos.shape).astype(np.float64)

    def set_masses(self, masses=1.0):
        """Simulation.py:232-276."""
        n = self._shape()[0]
        if isinstance(masses, float) or isinstance(masses, int):
            m = np.ones(shape=(n, 1), dtype=np.float64) * np.float64(masses)
        elif isinstance(masses, list):
            m = np.array(masses, dtype=np.float64).reshape(-1, 1)
        elif isinstance(masses, np.ndarray):
            m = masses.reshape(-1, 1)
        else:
            raise ValueError(f"Unknown type of masses: {type(masses)}")
        expected_shape = (n, 1)
        if m.shape != expected_shape:
            raise ValueError(f"Expected shape of masses: {expected_shape}, got: {m.shape}")
        self.masses = m
        if m.shape != self.particle_types.shape:
            self.particle_types = np.zeros(shape=m.shape, dtype=np.int32)

    def set_random_velocities(self, temperature: float = 1.0):
        """Simulation.py:278-290 (no net-momentum removal, as the reference)."""
        self.velocities = np.random.normal(loc=0.0, scale=np.sqrt(temperature / self.masses),
                                           size=self._shape()).astype(np.float64)

    def set_types(self, types=0):
        """Simulation.py:292-305."""
        n = self._shape()[0]
        if isinstance(types, int):
            t = np.ones(shape=(n, 1), dtype=np.int32) * np.int32(types)
        elif isinstance(types, list):
            t = np.array(types, dtype=np.int32).reshape(-1, 1)
        else:
            raise ValueError(f"Unknown type of types: {type(types)}")
        expected_shape = (n, 1)
        if t.shape != expected_shape:
            raise ValueError(f"Expected shape of types is {expected_shape} but got {t.shape}")
        self.particle_types = t

    def set_neighbor_list(self, skin: float = None, max_number_of_neighbors: int = None,
                          method_str: None | str = None) -> None:
        """Simulation.py:307-381: rebuild the list (cutoff = max sigma * r_cut
        + skin) with the GPU cell-list pipeline.  Neighbour sets equal those of
        the reference's double loop for either method string; rows come out
        ascending ("double loop") or in stencil order ("cell list")."""
        self.__dict__["_nl_host_count"] = self._nl_host_count + 1
        if skin is not None:
            self.neighbor_list_skin = np.float64(skin)
        if max_number_of_neighbors is not None:
            self.max_number_of_neighbors = np.int32(max_number_of_neighbors)
        if method_str is not None:
            if method_str not in self._KNOWN_CELL_LIST_METHODS:
                raise ValueError(f"Unknown neighbor list method: {method_str}. Try: {self._KNOWN_CELL_LIST_METHODS}.")
            self.neighbor_list_method_str = method_str
        if self.neighbor_list_method_str not in self._KNOWN_CELL_LIST_METHODS:
            raise ValueError(f"Unknown neighbor list method: {self.neighbor_list_method_str}")
        self.__dict__["_params_dirty"] = True
        self.__dict__["_nl_pending"] = True
        self.__dict__["_restore_xref"] = None
        if self._engine is not None:
            self._device()

    def update_neighbor_list(self, check=True) -> None:
        """Simulation.py:383-401."""
        if not check:
            self.set_neighbor_list()
            return
        eng = self._device()
        eng.update_neighbor_list()

    def set_pair_potential(self, pair_potential_str: str = "(1-r)**2", r_cut: float = 1.0, force_method=None,
                           energy_method=None) -> None:
        """Simulation.py:403-467."""
        if pair_potential_str in hardcoded_pair_potentials:
            entry = hardcoded_pair_potentials[pair_potential_str]
            self.__dict__["pair_potential_str"] = pair_potential_str
            self.__dict__["pair_potential"] = entry[1]
            self.__dict__["pair_force"] = entry[2]
            self.__dict__["pair_curvature"] = entry[3]
            self.__dict__["pair_potential_r_cut"] = entry[4]
            self.__dict__["_pot_dirty"] = True
            self.set_neighbor_list()
            return
        self.__dict__["pair_potential_str"] = pair_potential_str
        self.__dict__["pair_potential_r_cut"] = np.float64(r_cut)
        u, f, c = make_pair_potential(pair_potential_str=pair_potential_str, r_cut=r_cut)
        self.__dict__["pair_potential"] = u
        self.__dict__["pair_force"] = f
        self.__dict__["pair_curvature"] = c
        if force_method is not None:
            if force_method not in self._KNOWN_FORCE_METHODS:
                raise ValueError(f"Unknown force method: {force_method}. Try: {self._KNOWN_FORCE_METHODS}.")
            self.force_method_str = force_method
        if energy_method is not None:
            if energy_method not in self._KNOWN_ENERGY_METHODS:
                raise ValueError(f"Unknown energy method: {energy_method}. Try: {self._KNOWN_ENERGY_METHODS}.")
            self.energy_method_str = energy_method
        self.__dict__["_pot_dirty"] = True
        self.set_neighbor_list()

    def set_pair_potential_parameters(self, sigma: float = 1.0, epsilon: float = 1.0) -> None:
        """Simulation.py:469-525; callables are lowered to per-type tables."""
        if isinstance(sigma, float):
            self.__dict__["sigma_func"] = (lambda n, m, _s=sigma: np.float64(_s))
            self.__dict__["_sigma_const"] = float(sigma)
        elif callable(sigma):
            self.__dict__["sigma_func"] = sigma
            self.__dict__["_sigma_const"] = None
        else:
            raise ValueError(f"Unknown type of sigma: {type(sigma)}")
        if isinstance(epsilon, float):
            self.__dict__["epsilon_func"] = (lambda n, m, _e=epsilon: np.float64(_e))
            self.__dict__["_epsilon_const"] = float(epsilon)
        elif callable(epsilon):
            self.__dict__["epsilon_func"] = epsilon
            self.__dict__["_epsilon_const"] = None
        else:
            raise ValueError(f"Unknown type of epsilon: {type(epsilon)}")
        self.__dict__["_pot_dirty"] = True
        self.set_neighbor_list()

    def scale_box(self, scale_factor: float) -> None:
        """Simulation.py:527-543."""
        self.box_vectors = self.box_vectors * scale_factor
        self.positions = self.positions * scale_factor
        self.set_neighbor_list()

    def get_potential_energy(self) -> float:
        """Simulation.py:545-592 (every energy method string uses the GPU
        list kernel; the sums agree up to summation order)."""
        if self.energy_method_str not in self._KNOWN_ENERGY_METHODS:
            return None
        eng = self._device()
        return float(eng.compute(forces=False, energy=True)["energy"][0])

    def get_forces(self) -> np.ndarray:
        """Simulation.py:594-654."""
        if self.force_method_str not in self._KNOWN_FORCE_METHODS:
            raise ValueError(f"Unknown force method: {self.force_method_str}")
        eng = self._device()
        return eng.compute(forces=True)["forces"]

    def wrap_into_box(self) -> None:
        """Simulation.py:656-679 (host copy, in place)."""
        from .positions import wrap_into_box

        pos = self._get_field("positions")
        img = self._get_field("image_positions")
        wrap_into_box(pos.view(np.ndarray), img.view(np.ndarray), np.asarray(self.box_vectors, dtype=np.float64))
        self._dirty.update(("positions", "image_positions"))

    def _check_masses(self):
        bad = self.__dict__.get("_bad_fields", {})
        if bad:
            f, shp = next(iter(bad.items()))
            raise ValueError(f"{f} has shape {shp}, expected "
                             f"{(self.number_of_particles(), 1 if f == 'masses' else self.get_dimensions_of_space())}")

    def step(self, noise=None) -> None:
        """Simulation.py:681-704: skin check, (rebuild), forces, G-JF step,
        wrap -- one fused device sequence.  ``noise`` optionally supplies the
        standard normals [N, D] of integrator.py:30."""
        eng = self._device()
        self._check_masses()
        eng.step(1, "gjf", xi=noise)
        self._after_device_step()
        self.__dict__["number_of_steps"] = self.number_of_steps + 1

    def step_leap_frog(self) -> None:
        """Simulation.py:706-724."""
        eng = self._device()
        self._check_masses()
        eng.step(1, "leapfrog")
        self._after_device_step()
        self.__dict__["number_of_steps"] = self.number_of_steps + 1

    def run(self, steps: int = 1000) -> None:
        """Simulation.py:726-737: ``steps`` G-JF steps (captured in hipGraph
        blocks on the device)."""
        if steps <= 0:
            return
        eng = self._device()
        self._check_masses()
        eng.step(int(steps), "gjf")
        self._after_device_step()
        self.__dict__["number_of_steps"] = self.number_of_steps + int(steps)

    def set_integrator(self, time_step: float = None, target_temperature: float = None,
                       temperature_damping_time: float = None) -> None:
        """Simulation.py:739-759."""
        if time_step is not None:
            self.time_step = np.float64(time_step)
        if target_temperature is not None:
            self.temperature_target = np.float64(target_temperature)
        if temperature_damping_time is not None:
            self.temperature_damping_time = np.float64(temperature_damping_time)

    def set_kernel_tuning(self, threads_per_particle: int = 0, tile_particles: int = 0, fused: int = -1,
                          cell_scale: float = 1.0) -> None:
        """GPU kernel shape (no counterpart in the reference; set by
        dompap_amd.tools.autotune): lanes per particle (0: planner), home
        particles per tile (0: planner), fused force + integrate kernel (1),
        split (0) or planner's choice (-1), cells at least cell_scale x the
        list cutoff wide.  Results do not depend on it beyond summation
        order."""
        self.__dict__["_tuning"] = (int(threads_per_particle), int(tile_particles), int(fused))
        self.__dict__["_cell_scale"] = float(cell_scale)
        self.__dict__["_tuning_dirty"] = True

    def get_kernel_plan(self) -> dict:
        """The tile plan in use (lanes per particle, entry width, tiles, ...)."""
        return self._device().plan()

    def number_of_particles(self) -> int:
        return self._shape()[0]

    def get_density(self) -> float:
        return float(self.number_of_particles() / self.get_volume())

    def set_density(self, density: float = 1.0) -> None:
        """Simulation.py:787-802."""
        dimensions = self.box_vectors.shape[0]
        current_density = self.get_density()
        scale_factor = (current_density / density) ** (1 / dimensions)
        self.scale_box(scale_factor)

    def get_dimensions_of_space(self) -> int:
        return int(self.box_vectors.shape[0])

    def get_temperature(self) -> float:
        """Simulation.py:817-836."""
        m = self.masses[:, 0]
        v = np.asarray(self.velocities)
        v_squared = np.sum(v ** 2, axis=1)
        return float(np.sum(m * v_squared) / (self.get_dimensions_of_space() * self.number_of_particles()))

    def get_configurational_temperature(self) -> float:
        """Simulation.py:838-856."""
        forces = self.get_forces()
        numerator = np.mean(np.sum(forces ** 2, axis=1))
        denominator = np.mean(self.get_laplacian())
        return float(numerator / denominator)

    def get_radial_distribution_function(self, r_bins: np.ndarray):
        """Simulation.py:858-881: g(r) of all ordered pairs, self pairs
        included, over the minimum-image distances.  The pair counts are
        histogrammed on the GPU (dp_rdf_counts: the cell tiles of the current
        list when the last edge is within reach, else all pairs); the
        normalisation is positions.py:79-87's."""
        from .positions import normalise_radial_histogram

        r_bins = np.asarray(r_bins, dtype=np.float64)
        if r_bins.ndim != 1 or r_bins.size < 2:
            raise ValueError("r_bins must be a 1-D array of at least two bin edges")
        if not np.all(np.diff(r_bins) >= 0):
            raise ValueError("r_bins array must be sorted in ascending order.")
        eng = self._device()
        counts = eng.rdf_counts(r_bins).astype(np.int64)
        n = self.number_of_particles()
        return normalise_radial_histogram(counts, r_bins, n * n, np.asarray(self.box_vectors),
                                          self.get_dimensions_of_space())

    def get_number_of_particles(self) -> int:
        return self._shape()[0]

    def get_kinetic_energy(self) -> float:
        """Simulation.py:896-910."""
        v = np.asarray(self.velocities)
        m = np.asarray(self.masses)
        return float(0.5 * np.sum(m * v * v))

    def get_diameters(self) -> np.ndarray:
        """Simulation.py:912-928."""
        n = self.number_of_particles()
        diameters = np.ones(shape=(n, 1), dtype=np.float64)
        for i in range(n):
            diameters[i] = self.sigma_func(i, i)
        return diameters

    def get_time(self) -> float:
        return float(self.number_of_steps * self.time_step)

    def get_virial(self) -> float:
        """Simulation.py:946-967: W = (1/D) sum_{n<m} eps f(r/sigma) r, summed
        over the neighbour list on the GPU (every pair with a non-zero term is
        in the list)."""
        eng = self._device()
        return float(eng.compute(forces=False, virial=True)["virial"][0])

    def get_pressure(self) -> float:
        """Simulation.py:969-988."""
        V = self.get_volume()
        W = self.get_virial()
        P_c = W / V
        D = self.get_dimensions_of_space()
        v = np.asarray(self.velocities)
        m = np.asarray(self.masses)
        P_id = np.sum(v * v * m) / D / V
        return float(P_c + P_id)

    def get_volume(self) -> float:
        return float(np.prod(self.box_vectors))

    def set_particle_types(self, types) -> None:
        """Simulation.py:1003-1017."""
        self.particle_types = np.ones_like(np.asarray(self.masses), dtype=np.int32) * types

    def get_laplacian(self) -> np.ndarray:
        """Simulation.py:1205-1229."""
        eng = self._device()
        return eng.compute(forces=False, laplacian=True)["laplacian"]

    # ------------------------------------------------------------------
    # checkpoint (Simulation.py:1019-1203)
    # ------------------------------------------------------------------
    def to_disk(self, particle_data="simulation.csv", meta_data="simulation.toml") -> None:
        from .io import write_state

        write_state(self, particle_data, meta_data)

    def from_disk(self, particle_data="simulation.csv", meta_data="simulation.toml", verbose=False,
                  set_only_particle_data=False) -> dict:
        from .io import read_state

        return read_state(self, particle_data, meta_data, verbose, set_only_particle_data)

    def close(self) -> None:
        """Release the GPU context (it is re-created on the next device call)."""
        if self._engine is not None:
            for f in list(self._stale):
                self._get_field(f)
            self.__dict__["_nl_host_count"] = self._nl_updates()
            self._engine.close()
            self.__dict__["_engine"] = None
            self._dirty.update(_DEVICE_FIELDS + _HOST_FIELDS)
