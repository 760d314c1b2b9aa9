"""Thermodynamic-force table files: io::dumpThermoForce / io::restoreThermoForce
(io/DumpThermoForce.cpp:24-64, io/RestoreThermoForce.cpp:24-90, DumpProfile.cpp:22-48).

Format: the first line holds the bin centres, every further line the force of one atom
type over the bins, values separated by single spaces and written as a C++ stream writes a
double by default (six significant digits, `%g`).  Restoring takes the bin width from the
first two centres and builds the thermodynamic force on the given subdomain with it."""
import numpy as np

from . import ThermodynamicForce


def _line(values):
    return " ".join("%g" % v for v in values)


def dump_thermo_force(filename, tf, type_id=None):
    """all types (one line each), or only type `type_id`, after the grid line"""
    grid = tf.grid()
    force = tf.get_force()
    types = range(force.shape[1]) if type_id is None else [int(type_id)]
    with open(filename, "w") as f:
        f.write(_line(grid) + "\n")
        for t in types:
            f.write(_line(force[:, t]) + "\n")


def restore_thermo_force(ctx, filename, box, target_densities=(1.0,), modulations=(1.0,), enforce_symmetry=False,
                         use_periodicity=False, max_num_forces=10):
    """a ThermodynamicForce on `box` whose bin width is the distance of the file's first two
    bin centres, holding the file's forces (at most max_num_forces types)"""
    with open(filename) as f:
        lines = [ln.split() for ln in f.read().splitlines() if ln.strip()]
    if not lines or len(lines[0]) < 2:
        raise ValueError(f"{filename}: the grid line needs at least two bin centres")
    grid = [float(w) for w in lines[0]]
    rows = [[float(w) for w in ln] for ln in lines[1:]]
    if len(rows) > max_num_forces:
        raise ValueError(f"{filename}: {len(rows)} forces, at most {max_num_forces}")
    if any(len(r) != len(grid) for r in rows):
        raise ValueError(f"{filename}: every force line needs one value per bin")
    tf = ThermodynamicForce(ctx, list(target_densities), box, grid[1] - grid[0], list(modulations),
                            enforce_symmetry, use_periodicity)
    forces = np.array(rows, dtype=np.float64).T if rows else np.zeros((len(grid), 0))
    if forces.shape != (tf.num_bins, tf.num_types):
        raise ValueError(f"{filename}: {forces.shape[1]} forces over {forces.shape[0]} bins, the restored "
                         f"thermodynamic force has {tf.num_types} types over {tf.num_bins} bins")
    tf.set_force(forces)
    return tf
