"""On-disk state: particle CSV + TOML metadata in dompap's format
(Simulation.to_disk / from_disk, dompap/Simulation.py:1019-1203).

Columns ``ptype,mass,pos_d...,vel_d...,imgpos_d...,beta_d...``; metadata keys
and formatting as ``toml.dumps`` writes them (the reference's own files under
examples/ are read back by the tests).
"""
from __future__ import annotations

import os

import numpy as np

try:  # Python >= 3.11
    import tomllib as _toml_reader
except ModuleNotFoundError:  # pragma: no cover - Python 3.10
    import tomli as _toml_reader


def _toml_value(v):
    if isinstance(v, bool):
        return "true" if v else "false"
    if isinstance(v, (int, np.integer)):
        return str(int(v))
    if isinstance(v, (float, np.floating)):
        return repr(float(v))
    if isinstance(v, str):
        return '"' + v.replace("\\", "\\\\").replace('"', '\\"') + '"'
    if isinstance(v, (list, tuple)):
        return "[ " + " ".join(_toml_value(x) + "," for x in v) + "]"
    raise TypeError(f"cannot write {type(v)} to TOML")


def dumps(d: dict) -> str:
    """TOML text in the layout of ``toml.dumps`` for flat dicts."""
    return "".join(f"{k} = {_toml_value(v)}\n" for k, v in d.items())


def write_state(sim, particle_data="simulation.csv", meta_data="simulation.toml") -> None:
    dims = sim.get_dimensions_of_space()
    header = "ptype,mass"
    for prefix in ("pos", "vel", "imgpos", "beta"):
        for d in range(dims):
            header += f",{prefix}_{d}"
    types = np.asarray(sim.particle_types)
    masses = np.asarray(sim.masses)
    pos = np.asarray(sim.positions)
    vel = np.asarray(sim.velocities)
    img = np.asarray(sim.image_positions)
    bet = np.asarray(sim.betas)
    rows = []
    for n in range(sim.number_of_particles()):
        parts = [f"{types[n, 0]}", f"{masses[n, 0]}"]
        parts += [f"{pos[n, d]}" for d in range(dims)]
        parts += [f"{vel[n, d]}" for d in range(dims)]
        parts += [f"{img[n, d]}" for d in range(dims)]
        parts += [f"{bet[n, d]}" for d in range(dims)]
        rows.append(",".join(parts))
    with open(particle_data, "w") as f:
        print(header + "".join("\n" + r for r in rows), file=f)
    meta = {
        "box_vectors": np.asarray(sim.box_vectors).tolist(),
        "dimensions_of_space": sim.get_dimensions_of_space(),
        "number_of_particles": sim.number_of_particles(),
        "temperature": sim.get_temperature(),
        "potential_energy": sim.get_potential_energy(),
        "kinetic_energy": sim.get_kinetic_energy(),
        "pressure": sim.get_pressure(),
        "virial": sim.get_virial(),
        "density": sim.get_density(),
        "volume": sim.get_volume(),
        "pair_potential_str": sim.pair_potential_str,
        "pair_potential_r_cut": float(sim.pair_potential_r_cut),
        "neighbor_list_skin": float(sim.neighbor_list_skin),
        "max_number_of_neighbors": int(sim.max_number_of_neighbors),
        "neighbor_list_method_str": sim.neighbor_list_method_str,
        "energy_method_str": sim.energy_method_str,
        "force_method_str": sim.force_method_str,
        "time_step": float(sim.time_step),
        "temperature_target": float(sim.temperature_target),
        "temperature_damping_time": float(sim.temperature_damping_time),
        "number_of_steps": int(sim.number_of_steps),
        "number_of_neighbor_list_updates": int(sim.number_of_neighbor_list_updates),
    }
    with open(meta_data, "w") as f:
        print(dumps(meta), file=f)


def read_meta(meta_data: str) -> dict:
    with open(meta_data, "rb") as f:
        return _toml_reader.load(f)


def read_particles(particle_data: str, number_of_particles: int, dims: int):
    """Parse the particle CSV (Simulation.py:1145-1171)."""
    types = np.zeros((number_of_particles, 1), dtype=np.int32)
    pos = np.zeros((number_of_particles, dims), dtype=np.float64)
    vel = np.zeros((number_of_particles, dims), dtype=np.float64)
    img = np.zeros((number_of_particles, dims), dtype=np.int32)
    bet = np.zeros((number_of_particles, dims), dtype=np.float64)
    mass = np.zeros((number_of_particles, 1), dtype=np.float64)
    with open(particle_data) as f:
        lines = f.readlines()
    col_names = [c.strip() for c in lines[0].split(",")]
    for i, line in enumerate(lines[1:]):
        if not line.strip():
            continue
        cols = line.split(",")
        for value, name in zip(cols, col_names):
            if "_" in name:
                kind, d = name.split("_")
                d = int(d)
                if kind == "pos":
                    pos[i, d] = float(value)
                elif kind == "vel":
                    vel[i, d] = float(value)
                elif kind == "imgpos":
                    img[i, d] = int(value)
                elif kind == "beta":
                    bet[i, d] = float(value)
                else:
                    raise ValueError(f"Unknown column name: {kind}")
            else:
                if name == "ptype":
                    types[i, 0] = int(value)
                elif name == "mass":
                    mass[i, 0] = float(value)
                else:
                    raise ValueError(f"Unknown column name: {name}")
    return types, mass, pos, vel, img, bet


def read_state(sim, particle_data="simulation.csv", meta_data="simulation.toml", verbose=False,
               set_only_particle_data=False) -> dict:
    if not os.path.isfile(particle_data):
        raise FileNotFoundError(f"File not found: {particle_data}")
    if not os.path.isfile(meta_data):
        raise FileNotFoundError(f"File not found: {meta_data}")
    meta = read_meta(meta_data)
    box = np.array(meta["box_vectors"], dtype=np.float64)
    dims = box.shape[0]
    n = int(meta["number_of_particles"])
    if verbose:
        print("    Old values")
        print(f"Dimensions of space: {sim.get_dimensions_of_space()}")
        print(f"number_of_particles: {sim.number_of_particles()}")
        print(f"Box vectors: {sim.box_vectors}")
        print("    New values")
        print(f"Dimensions of space: {dims}")
        print(f"Box vectors: {box}")
        print(f"number_of_particles: {n}")
    types, mass, pos, vel, img, bet = read_particles(particle_data, n, dims)
    sim.box_vectors = box
    sim.particle_types = types
    sim.positions = pos
    sim.velocities = vel
    sim.image_positions = img
    sim.betas = bet
    sim.masses = mass
    if not set_only_particle_data:
        sim.set_pair_potential(pair_potential_str=meta["pair_potential_str"], r_cut=meta["pair_potential_r_cut"],
                               force_method=meta["force_method_str"], energy_method=meta["energy_method_str"])
        sim.set_neighbor_list(skin=meta["neighbor_list_skin"], max_number_of_neighbors=meta["max_number_of_neighbors"],
                              method_str=meta["neighbor_list_method_str"])
        sim.set_integrator(time_step=meta["time_step"], target_temperature=meta["temperature_target"],
                           temperature_damping_time=meta["temperature_damping_time"])
        sim.__dict__["number_of_steps"] = meta["number_of_steps"]
        sim.number_of_neighbor_list_updates = meta["number_of_neighbor_list_updates"]
    return meta
