"""Files of the reference's io module on the hot path: the thermodynamic-force tables
(io::dumpThermoForce / io::restoreThermoForce, io/DumpThermoForce.cpp:24-64,
io/RestoreThermoForce.cpp:24-90, DumpProfile.cpp:22-48) and GROMACS coordinate files
(io::restoreGRO, io/RestoreGRO.cpp:24-147).

Format: the first line holds the bin centres, every further line the force of one atom
type over the bins, values separated by single spaces and written as a C++ stream writes a
double by default (six significant digits, `%g`).  Restoring takes the bin width from the
first two centres and builds the thermodynamic force on the given subdomain with it."""
import re

import numpy as np

from . import Atoms, ThermodynamicForce, make_box


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


_INT = re.compile(r"[+-]?\d+")
_FLOAT = re.compile(r"[+-]?(\d+\.?\d*|\.\d+)([eE][+-]?\d+)?")


def _scan(line, fields):
    """C sscanf over width-limited conversions, as the reference scans a GRO atom line: each
    (width, kind) skips leading white space, then reads at most `width` characters of an
    integer ('d'), a floating value ('f') or non-space characters ('s').  Returns the values
    converted before the first failure (sscanf's count is their number)."""
    out, pos = [], 0
    for width, kind in fields:
        while pos < len(line) and line[pos].isspace():
            pos += 1
        chunk = line[pos:pos + width]
        if kind == "s":
            m = re.match(r"\S+", chunk)
        else:
            m = (_INT if kind == "d" else _FLOAT).match(chunk)
        if not m or not m.group(0):
            break
        tok = m.group(0)
        out.append(tok if kind == "s" else (int(tok) if kind == "d" else float(tok)))
        pos += len(tok)
    return out


_GRO_ATOM = [(5, "d"), (5, "s"), (5, "s"), (5, "d")] + [(8, "f")] * 6


def read_gro(filename):
    """the contents of a GROMACS coordinate file as io::restoreGRO reads them: positions,
    velocities (zero for lines without them) and the box edges of the last line"""
    with open(filename) as f:
        lines = f.read().splitlines()
    if len(lines) < 2:
        raise ValueError(f"{filename}: no atom count line")
    n = int(_scan(lines[1], [(len(lines[1]) + 1, "d")])[0])
    if n <= 0:
        raise ValueError(f"{filename}: invalid atom count {n}")
    if len(lines) < n + 3:
        raise ValueError(f"{filename}: {n} atoms announced, {len(lines) - 3} lines present")
    pos = np.zeros((n, 3))
    vel = np.zeros((n, 3))
    for i in range(n):
        vals = _scan(lines[2 + i], _GRO_ATOM)
        if len(vals) not in (7, 10):
            raise ValueError(f"Invalid GRO atom line in file {filename}: expected 7 fields for coordinates (or 10 "
                             f"fields including velocities), but parsed {len(vals)} fields from line: "
                             f"\"{lines[2 + i]}\"")
        pos[i] = vals[4:7]
        if len(vals) == 10:
            vel[i] = vals[7:10]
    edges = np.array([float(w) for w in lines[2 + n].split()[:3]])
    return pos, vel, edges


def restore_gro(ctx, filename, min_corner=(0.0, 0.0, 0.0), ghost_layer=0.0):
    """io::restoreGRO: the file's atoms as local atoms (mass and relative mass 1, type 0,
    force 0) and the subdomain from `min_corner` over the file's box edges with the given
    ghost layer thickness.  Returns (Atoms, box, edges)."""
    pos, vel, edges = read_gro(filename)
    lo = np.asarray(min_corner, dtype=np.float64)
    atoms = Atoms.from_arrays(ctx, pos, vel=vel, force=np.zeros_like(pos))
    return atoms, make_box(lo, lo + edges, ghost_layer), edges
