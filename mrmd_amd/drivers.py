"""Time loops of the reference's example drivers, written against the action API.

run_lj mirrors examples/02_LennardJones_NVE/02_LennardJones_NVE.cpp:135-186 (velocity
Verlet) and, with langevin=True, the pre-force integrator of
examples/01_LennardJones_NVT/01_LennardJones_NVT.cpp:139-214; permute=True adds the
LinkedCellList + Atoms::permute of tests/NVT/NVT.cpp:138-144 on every rebuild.
"""
import numpy as np

from . import (FORCE, POS, VEL, Atoms, GhostLayer, HalfVerletList, LennardJones, LinkedCellList,
               VelocityVerlet, VelocityVerletLangevinThermostat, make_box)


class LJSystem:
    def __init__(self, ctx, pos, vel, L, rc=2.5, skin=0.3, sigma=1.0, epsilon=1.0, cap=0.0, shifted=True,
                 cell_ratio=1.0, max_neigh=0, langevin=False, gamma=1.0, temperature=1.0, seed=1234,
                 permute=False, types=None, num_types=1, type_stride=1):
        self.ctx = ctx
        self.L = L
        self.rc, self.skin = rc, skin
        self.rn = rc + skin
        self.box = make_box([0, 0, 0], [L] * 3, self.rn)
        self.atoms = Atoms.from_arrays(ctx, pos, vel=vel, type=types)
        self.ghosts = GhostLayer(ctx)
        self.vl = HalfVerletList(ctx)
        self.lj = LennardJones(ctx, rc, sigma, epsilon, cap, num_types, shifted, type_stride)
        self.cell_ratio, self.max_neigh = cell_ratio, max_neigh
        self.langevin = VelocityVerletLangevinThermostat(gamma, temperature, seed) if langevin else None
        self.permute = permute
        self.max_disp = np.finfo(np.float64).max
        self.rebuilds = 0
        self.step_count = 0

    def rebuild(self):
        a = self.atoms
        self.ghosts.exchange_real_atoms(a, self.box)
        if self.permute:
            cl = LinkedCellList(self.ctx, a, 0, a.nlocal, [self.rn] * 3, [0, 0, 0], [self.L] * 3)
            a.permute(cl)
        self.ghosts.create_ghost_atoms(a, self.box)
        lo, hi = -self.rn, self.L + self.rn
        self.vl.build(a, 0, a.nlocal, self.rn, self.cell_ratio, [lo] * 3, [hi] * 3, self.max_neigh)
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

    def run(self, nsteps, dt, energy_every=0):
        rec = []
        for s in range(nsteps):
            self.step(dt)
            if energy_every and s % energy_every == 0:
                e, w = self.lj.energy_virial()
                rec.append((e, self.atoms.kinetic_energy(), w))
        return np.array(rec)

    def local(self, field):
        return self.atoms.download(field, self.atoms.nlocal)
