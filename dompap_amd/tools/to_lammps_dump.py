"""LAMMPS dump text of a Simulation's state (tools/to_lammps_dump.py:4-48).

The state and the forces come from the device (``Simulation`` attributes
download on access, ``get_forces`` runs the pair kernel); the text is built
once with ``str.join`` rather than by repeated concatenation.  The layout is
the reference's, including its quirks: only 3-D systems are written (the 2-D
branch of the reference ends in the ``else`` of its 3-D test and raises),
the ATOMS header and every atom line are joined by newlines with no newline
at the end, and numbers are printed with Python's ``str`` of a float.
"""
from __future__ import annotations

import numpy as np


def to_lammps_dump(sim) -> str:
    """Return positions, box and per-atom state as LAMMPS dump text."""
    number_of_steps = sim.number_of_steps
    box_vectors = sim.box_vectors
    particle_types = np.asarray(sim.particle_types).reshape(-1)
    masses = np.asarray(sim.masses).reshape(-1)
    diameters = np.asarray(sim.get_diameters()).reshape(-1)
    positions = np.asarray(sim.positions)
    image_positions = np.asarray(sim.image_positions)
    velocities = np.asarray(sim.velocities)
    forces = sim.get_forces()

    number_of_atoms = positions.shape[0]
    dimension_of_space = positions.shape[1]
    head = ['ITEM: TIMESTEP', f'{number_of_steps:d}', 'ITEM: NUMBER OF ATOMS', f'{number_of_atoms:d}']
    if dimension_of_space != 3:
        raise ValueError('Dimension of space must be 2 or 3')
    head.append('ITEM: BOX BOUNDS pp pp pp')
    head += [f'0.0 {box_vectors[d]}' for d in range(3)]
    head.append('ITEM: ATOMS id type diameter mass x y z ix iy iz vx vy vz fx fy fz')
    x = positions.tolist()
    im = image_positions.astype(np.int64).tolist()
    v = velocities.tolist()
    f = np.asarray(forces).tolist()
    t = particle_types.astype(np.int64).tolist()
    dia = diameters.astype(np.float64).tolist()
    m = masses.astype(np.float64).tolist()
    rows = [f'{n} {t[n]} {dia[n]} {m[n]} '
            f'{x[n][0]} {x[n][1]} {x[n][2]} '
            f'{im[n][0]} {im[n][1]} {im[n][2]} '
            f'{v[n][0]} {v[n][1]} {v[n][2]} '
            f'{f[n][0]} {f[n][1]} {f[n][2]}' for n in range(number_of_atoms)]
    return '\n'.join(head + rows)
