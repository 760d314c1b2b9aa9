"""Time loops of the reference's example drivers, written against the action API.

LJSystem.step mirrors examples/02_LennardJones_NVE/02_LennardJones_NVE.cpp:135-186
(velocity Verlet) and, with langevin=True, the pre-force integrator of
examples/01_LennardJones_NVT/01_LennardJones_NVT.cpp:139-214; permute=True adds the
LinkedCellList + Atoms::permute of tests/NVT/NVT.cpp:138-144 on every rebuild.

AdResSSystem.step is the LJ/ideal-gas AdResS step composed from the reference's actions in
the order their contracts force (molecule forces reach ghost atoms before the ghost fold;
lambda is current before LJ_IdealGas::run): Langevin pre-force integration,
UpdateMolecules, multi-resolution ghost exchange / Verlet list over molecule centres on a
rebuild (else a ghost refresh), UpdateMolecules, LJ_IdealGas::run,
ContributeMoleculeForceToAtoms, ghost fold, ThermodynamicForce sample / periodic update /
apply, Langevin post-force integration (action/LJ_IdealGas.cpp:227-260,
action/ContributeMoleculeForceToAtoms.cpp:23-48, action/ThermodynamicForce.hpp:98-152).
"""
import ctypes as C

import numpy as np

from . import _L
from .capi import AdressLoop, MdLoop
from . import (FORCE, POS, PRED_SLAB_HY, VEL, Atoms, GhostLayer, HalfVerletList, LennardJones, LinkedCellList,
               LJ_IdealGas, Molecules, Slab, ThermodynamicForce, VelocityVerlet, VelocityVerletLangevinThermostat,
               make_box, pred1)


def _edges(L):
    e = np.broadcast_to(np.asarray(L, dtype=np.float64), (3,)).copy()
    return e


def list_grid(L, rn, cell_ratio=1.0):
    """Binning grid of the Verlet list: the reference's ghost corners [-r_n, L + r_n] with
    cells of cell_ratio r_n, except that for cell_ratio 1 and a box whose three edges give
    the cell list the same cell edge c = L / floor(L / r_n), the grid is [-c, L + c] in
    cells of c: the cell list's cells plus one layer for the ghosts, so that the list's
    slot order follows the order in which Atoms::permute leaves the atoms.  Any grid of
    cells no smaller than r_n that covers the ghosts gives the same neighbour sets.
    Returns (grid_min, grid_max, cell_ratio)."""
    L = np.asarray(L, dtype=np.float64)
    if cell_ratio == 1.0:
        nc = np.floor(L / rn)
        c = L / nc
        if np.all(nc >= 1) and np.all(c == c[0]):
            delta = c[0] * (1.0 - 1e-9)  # floor((L + 2c) / delta) is then nc + 2 beyond doubt
            lo, hi = -c, L + c
            if delta >= rn and np.all(np.floor((hi - lo) * (1.0 / delta)) == nc + 2):
                return lo, hi, delta / rn
    return np.full(3, -rn), L + rn, cell_ratio


class LJSystem:
    def __init__(self, ctx, pos, vel, L, rc=2.5, skin=0.3, sigma=1.0, epsilon=1.0, cap=0.0, shifted=True,
                 cell_ratio=1.0, max_neigh=0, langevin=False, gamma=1.0, temperature=1.0, seed=1234,
                 permute=False, types=None, num_types=1, type_stride=1):
        self.ctx = ctx
        self.L = _edges(L)
        self.rc, self.skin = rc, skin
        self.rn = rc + skin
        self.box = make_box([0, 0, 0], self.L, self.rn)
        self.atoms = Atoms.from_arrays(ctx, pos, vel=vel, type=types)
        self.ghosts = GhostLayer(ctx)
        self.vl = HalfVerletList(ctx)
        self.lj = LennardJones(ctx, rc, sigma, epsilon, cap, num_types, shifted, type_stride)
        self.cell_ratio, self.max_neigh = cell_ratio, max_neigh
        self.grid_min, self.grid_max, self.list_ratio = list_grid(self.L, self.rn, cell_ratio)
        self.langevin = VelocityVerletLangevinThermostat(gamma, temperature, seed) if langevin else None
        self.permute = permute
        self.cells = None
        self.max_disp = np.finfo(np.float64).max
        self.rebuilds = 0
        self.step_count = 0

    def rebuild(self):
        a = self.atoms
        self.ghosts.exchange_real_atoms(a, self.box)
        if self.permute:
            if self.cells is None:
                self.cells = LinkedCellList(self.ctx, a, 0, a.nlocal, [self.rn] * 3, [0, 0, 0], self.L)
            else:
                self.cells.build(a, 0, a.nlocal, [self.rn] * 3, [0, 0, 0], self.L)
            a.permute(self.cells)
        self.ghosts.create_ghost_atoms(a, self.box)
        self.vl.build(a, 0, a.nlocal, self.rn, self.list_ratio, self.grid_min, self.grid_max, self.max_neigh)
        self.rebuilds += 1

    def step(self, dt):
        a = self.atoms
        if self.langevin is not None:
            self.max_disp += self.langevin.pre_force_integrate(a, dt)
        else:
            self.max_disp += VelocityVerlet.pre_force_integrate(a, dt)
        if self.max_disp >= self.skin * 0.5:
            self.max_disp = 0.0
            self.rebuild()
        else:
            self.ghosts.update_ghost_atoms(a, self.box)
        a.set_force(0.0)
        self.lj.apply(a, self.vl)
        self.ghosts.contribute_back_ghost_to_real(a)
        VelocityVerlet.post_force_integrate(a, dt)
        self.step_count += 1

    def run(self, nsteps, dt, energy_every=0, native=True):
        """nsteps steps.  Without energy records they run in the library's native loop
        (mrmd_md_run: the calls of step(), in its order, made from C++), otherwise step by
        step; both give the same results."""
        if energy_every or not native:
            rec = []
            for s in range(nsteps):
                self.step(dt)
                if energy_every and s % energy_every == 0:
                    e, w = self.lj.energy_virial()
                    rec.append((e, self.atoms.kinetic_energy(), w))
            return np.array(rec)
        a = self.atoms
        if self.permute and self.cells is None:
            self.cells = LinkedCellList(self.ctx, a, 0, a.nlocal, [self.rn] * 3, [0, 0, 0], self.L)
        lang = self.langevin
        loop = MdLoop(dt=dt, neighbor_radius=self.rn, skin=self.skin, cell_ratio=self.list_ratio,
                      grid_min=(C.c_double * 3)(*self.grid_min), grid_max=(C.c_double * 3)(*self.grid_max),
                      max_neigh=int(self.max_neigh), langevin=1 if lang is not None else 0,
                      zeta=lang.zeta if lang else 0.0, temperature=lang.temperature if lang else 0.0,
                      seed=lang.seed if lang else 0, call=lang.call if lang else 0, max_disp=self.max_disp,
                      rebuilds=self.rebuilds, steps=self.step_count)
        self.ctx.check(_L.mrmd_md_run(a.h, self.ghosts.h, self.vl.h, self.lj.h,
                                      self.cells.h if self.permute else None, C.byref(self.box), C.byref(loop),
                                      int(nsteps)))
        self.max_disp, self.rebuilds, self.step_count = loop.max_disp, loop.rebuilds, loop.steps
        if lang is not None:
            lang.call = loop.call
        return np.zeros((0, 3))

    def local(self, field):
        return self.atoms.download(field, self.atoms.nlocal)


class AdResSSystem:
    """monatomic LJ / ideal-gas AdResS slab: one molecule per atom (createMoleculeForEachAtom),
    slab weighting along x centred in the box, thermodynamic force on x-bins"""

    def __init__(self, ctx, pos, vel, L, atomistic_diameter, hybrid_diameter, nu=1, rc=2.5, skin=0.3,
                 sigma=1.0, epsilon=1.0, cap=0.0, shifted=True, gamma=1.0, temperature=1.0, seed=1234,
                 target_density=None, bin_width=1.0, tf_every=100, tf_sigma=2.0, tf_range=3.0,
                 tf_symmetric=True, tf_periodic=True, modulation=1.0):
        self.ctx = ctx
        self.L = _edges(L)
        self.rc, self.skin = rc, skin
        self.rn = rc + skin
        self.box = make_box([0, 0, 0], self.L, self.rn)
        self.atoms = Atoms.from_arrays(ctx, pos, vel=vel)
        self.mols = Molecules.for_each_atom(ctx, self.atoms)
        self.slab = Slab(0.5 * self.L, atomistic_diameter, hybrid_diameter, nu)
        self.ghosts = GhostLayer(ctx)
        self.vl = HalfVerletList(ctx)
        self.ljig = LJ_IdealGas(ctx, [cap], [rc], [sigma], [epsilon], shifted, 1)
        rho = target_density if target_density is not None else pos.shape[0] / float(np.prod(self.L))
        self.tf = ThermodynamicForce(ctx, [rho], self.box, bin_width, [modulation], tf_symmetric, tf_periodic)
        self.tf_every, self.tf_sigma, self.tf_range = tf_every, tf_sigma, tf_range
        # update_if over the hybrid region: the thermodynamic force acts where lambda varies
        self.tf_pred = pred1(kind=PRED_SLAB_HY, axis=0, a=0.5 * self.L[0], b=0.5 * atomistic_diameter,
                             c=hybrid_diameter)
        self.langevin = VelocityVerletLangevinThermostat(gamma, temperature, seed)
        self.max_disp = np.finfo(np.float64).max
        self.rebuilds = 0
        self.step_count = 0
        self.energy = 0.0

    def rebuild(self):
        m, a = self.mols, self.atoms
        m.update(a, None)  # molecule centres decide the exchange
        self.ghosts.multires_exchange_real_atoms(m, a, self.box)
        self.ghosts.multires_create_ghost_atoms(m, a, self.box)
        m.update(a, None)
        self.vl.build(m, 0, m.nlocal, self.rn, 1.0, np.full(3, -self.rn), self.L + self.rn, molecules=True)
        self.rebuilds += 1

    def step(self, dt):
        m, a = self.mols, self.atoms
        self.max_disp += self.langevin.pre_force_integrate(a, dt)
        if self.max_disp >= self.skin * 0.5:
            self.max_disp = 0.0
            self.rebuild()
        else:
            self.ghosts.update_ghost_atoms(a, self.box)
        m.update(a, self.slab)
        a.set_force(0.0)
        m.set_force(0.0)
        self.energy = self.ljig.run(m, self.vl, a)
        m.contribute_force(a)
        self.ghosts.contribute_back_ghost_to_real(a)
        self.tf.sample(a)
        self.step_count += 1
        if self.tf_every and self.step_count % self.tf_every == 0:
            self.tf.update(self.tf_sigma, self.tf_range, self.tf_pred)
        self.tf.apply(a)
        self.langevin.post_force_integrate(a, dt)

    def run(self, nsteps, dt, native=True):
        """nsteps steps: in the library's native loop (mrmd_adress_md_run: the calls of
        step(), in its order, made from C++; the energy is read back on the last step only)
        or step by step; both give the same results"""
        if not native:
            for _ in range(nsteps):
                self.step(dt)
            return
        if nsteps <= 0:
            return
        lang = self.langevin
        loop = AdressLoop(dt=dt, neighbor_radius=self.rn, skin=self.skin, zeta=lang.zeta,
                          temperature=lang.temperature, seed=lang.seed, tf_every=int(self.tf_every or 0),
                          tf_sigma=self.tf_sigma, tf_range=self.tf_range, tf_pred=self.tf_pred, call=lang.call,
                          max_disp=self.max_disp, rebuilds=self.rebuilds, steps=self.step_count)
        self.ctx.check(_L.mrmd_adress_md_run(self.atoms.h, self.mols.h, self.ghosts.h, self.vl.h, self.ljig.h,
                                             self.tf.h, C.byref(self.slab.params), C.byref(self.box), C.byref(loop),
                                             int(nsteps)))
        self.max_disp, self.rebuilds, self.step_count = loop.max_disp, loop.rebuilds, loop.steps
        lang.call = loop.call
        self.energy = loop.energy

    def local(self, field):
        return self.atoms.download(field, self.atoms.nlocal)
