"""Checkpoints of one system (LJSystem, AdResSSystem).

The particle datasets and the subdomain are those of the reference's H5MD dump and restore
(io/DumpH5MD.hpp:30-75, io/DumpH5MD.cpp:224-278, io/RestoreH5MD.cpp:90-190): position,
velocity, force, type, mass, charge and relativeMass of the local atoms, the box edges and
the subdomain's minCorner / maxCorner / ghostLayerThickness, under H5MD's paths
(particles/<group>/<dataset>/value, particles/<group>/box/...).  HDF5 is not part of this
image, so the container is a NumPy .npz archive with those paths as member names.

A checkpoint also holds what a bit-identical resume needs beyond H5MD: the step count, the
Langevin noise counter and seed, the rebuild count and, for AdResS, the LJ_IdealGas run
counter and compensation-energy histograms (LJ_IdealGas.cpp:21-50, 241-257) and the
thermodynamic force with its accumulated density profile (ThermodynamicForce.cpp:74-86).

Restored atoms are local only, as in the reference's restore (numGhostAtoms = 0,
RestoreH5MD.cpp:181-183); the system rebuilds its ghosts and Verlet list at its next
step.  A dump ends the running system's list epoch too (its next step rebuilds), so a run
that goes on after a dump and a run restored from it take the same steps: bit-identical
for LJSystem; the AdResS step adds ghost-partner forces with atomics, so there the two
agree to rounding.
"""
import numpy as np

from . import CHARGE, FORCE, MASS, POS, RELMASS, TYPE, VEL, Molecules

FORMAT = "mrmd_amd-checkpoint-1"
DATASETS = (("position", POS), ("velocity", VEL), ("force", FORCE), ("type", TYPE), ("mass", MASS),
            ("charge", CHARGE), ("relativeMass", RELMASS))


def write_frame(path, frame):
    """frame: {member path: array}; written uncompressed, every member exactly"""
    np.savez(path, **{k: np.asarray(v) for k, v in frame.items()})


def read_frame(path):
    with np.load(path, allow_pickle=False) as f:
        frame = {k: f[k] for k in f.files}
    if str(frame.get("format", "")) != FORMAT:
        raise ValueError(f"{path} is not a {FORMAT} file")
    return frame


def _box_members(system, group):
    b = system.box
    lo = np.array([b.min[d] for d in range(3)])
    hi = np.array([b.max[d] for d in range(3)])
    ghost = np.array([b.ghost[d] for d in range(3)])
    p = f"particles/{group}/box"
    return {f"{p}/edges/value": hi - lo, f"{p}/minCorner": lo, f"{p}/maxCorner": hi,
            f"{p}/ghostLayerThickness": ghost, f"{p}/boundary": np.array(["periodic"] * 3)}


def frame_of(system, group="atoms", author="xxx"):
    """the checkpoint frame of a system (downloads; no state changes)"""
    a = system.atoms
    n = a.nlocal
    frame = {"format": np.array(FORMAT), "h5md/author/name": np.array(author),
             "h5md/creator/name": np.array("mrmd_amd")}
    frame.update(_box_members(system, group))
    for name, field in DATASETS:
        frame[f"particles/{group}/{name}/value"] = a.download(field, n)
        frame[f"particles/{group}/{name}/step"] = np.array(system.step_count, np.int64)
    lang = system.langevin
    frame["integrator/step"] = np.array(system.step_count, np.int64)
    frame["integrator/rebuilds"] = np.array(system.rebuilds, np.int64)
    if lang is not None:
        frame["langevin/seed"] = np.array(lang.seed, np.uint64)
        frame["langevin/call"] = np.array(lang.call, np.int64)
        frame["langevin/zeta_temperature"] = np.array([lang.zeta, lang.temperature])
    if hasattr(system, "ljig"):
        ig = system.ljig
        frame["lj_idealgas/run_counter"] = np.array(ig.run_counter, np.int64)
        for k in range(3):
            frame[f"lj_idealgas/histogram{k}"] = ig.histogram(k)
        tf = system.tf
        frame["thermodynamic_force/force"] = tf.get_force()
        frame["thermodynamic_force/density"] = tf.get_density_profile()
        frame["thermodynamic_force/samples"] = np.array(tf.samples, np.int64)
        frame["lj_idealgas/energy"] = np.array(system.energy)
    return frame


def dump(path, system, group="atoms", author="xxx"):
    """write a checkpoint of `system`; its next step rebuilds the Verlet list (see above)"""
    write_frame(path, frame_of(system, group, author))
    system.max_disp = np.finfo(np.float64).max


def restore(path, system, group="atoms"):
    """load a checkpoint into `system`, which was constructed with the parameters of the
    run that wrote it (cut-offs, potential, thermostat, AdResS geometry); the atom count
    may differ from the constructor's arrays"""
    frame = read_frame(path)
    lo = frame[f"particles/{group}/box/minCorner"]
    hi = frame[f"particles/{group}/box/maxCorner"]
    b = system.box
    if not (np.array_equal(lo, [b.min[d] for d in range(3)]) and np.array_equal(hi, [b.max[d] for d in range(3)])):
        raise ValueError(f"checkpoint box {lo} - {hi} differs from the system's")
    pos = frame[f"particles/{group}/position/value"]
    n = pos.shape[0]
    a = system.atoms
    a.set_counts(0, 0)
    if a.size < n:
        a.resize(n)
    for name, field in DATASETS:
        a.upload(field, frame[f"particles/{group}/{name}/value"])
    a.set_counts(n, 0)
    system.step_count = int(frame["integrator/step"])
    system.rebuilds = int(frame["integrator/rebuilds"])
    system.max_disp = np.finfo(np.float64).max
    if system.langevin is not None:
        if "langevin/call" not in frame:
            raise ValueError("checkpoint of a run without Langevin noise counter")
        system.langevin.seed = int(frame["langevin/seed"])
        system.langevin.call = int(frame["langevin/call"])
    if hasattr(system, "ljig"):
        if "lj_idealgas/run_counter" not in frame:
            raise ValueError("checkpoint holds no AdResS state")
        system.mols = Molecules.for_each_atom(system.ctx, a)
        ig = system.ljig
        ig.run_counter = int(frame["lj_idealgas/run_counter"])
        for k in range(3):
            ig.set_histogram(k, frame[f"lj_idealgas/histogram{k}"])
        tf = system.tf
        tf.set_force(frame["thermodynamic_force/force"])
        tf.set_density_profile(frame["thermodynamic_force/density"], int(frame["thermodynamic_force/samples"]))
        system.energy = float(frame["lj_idealgas/energy"])
    return frame
