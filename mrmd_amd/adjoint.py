"""Force matching with the adjoint LJ kernel: a torch.autograd.Function whose forward is
the device force pass and whose backward is the device adjoint (dL/dx and dL/d(eps, sigma)
per type pair, pair terms recomputed), plus Lorentz-Berthelot mixing and an Adam fit.

There is no reference counterpart; the forward is action::LennardJones with the
type-correct pair index t_i * n + t_j (the convention of LJ_IdealGas.cpp:139).
"""
import numpy as np

from . import (FORCE, POS, Atoms, GhostLayer, HalfVerletList, LennardJones, VelocityVerlet,
               VelocityVerletLangevinThermostat, make_box)


def lorentz_berthelot(eps_types, sigma_types):
    """per-type (eps_a, sigma_a) -> row-major pair tables eps_ab = sqrt(eps_a eps_b),
    sigma_ab = (sigma_a + sigma_b) / 2"""
    import torch
    eps_tab = torch.sqrt(eps_types[:, None] * eps_types[None, :])
    sig_tab = 0.5 * (sigma_types[:, None] + sigma_types[None, :])
    return eps_tab.reshape(-1), sig_tab.reshape(-1)


class Snapshot:
    """one configuration resident on the device: atoms, ghosts, Verlet list, LJ table"""

    def __init__(self, ctx, pos, types, L, num_types, eps, sigma, rc=2.5, skin=0.3):
        self.ctx = ctx
        self.n = pos.shape[0]
        self.L = L
        rn = rc + skin
        self.box = make_box([0, 0, 0], [L] * 3, rn)
        self.atoms = Atoms.from_arrays(ctx, pos, type=types)
        self.ghosts = GhostLayer(ctx)
        self.ghosts.exchange_real_atoms(self.atoms, self.box)
        self.ghosts.create_ghost_atoms(self.atoms, self.box)
        self.vl = HalfVerletList(ctx)
        self.vl.build(self.atoms, 0, self.atoms.nlocal, rn, 1.0, [-rn] * 3, [L + rn] * 3)
        self.lj = LennardJones(ctx, rc, sigma, eps, 0.0, num_types, shifted=True, type_stride=num_types)
        self.num_types = num_types

    def set_params(self, eps, sigma):
        self.lj.set_params(np.asarray(sigma, dtype=np.float64), np.asarray(eps, dtype=np.float64))

    def forces(self):
        self.atoms.set_force(0.0)
        self.lj.apply(self.atoms, self.vl)
        self.ghosts.contribute_back_ghost_to_real(self.atoms)
        return self.atoms.download(FORCE, self.atoms.nlocal)

    def adjoint(self, dLdF):
        return self.lj.adjoint(self.atoms, self.vl, dLdF, ghost_layer=self.ghosts)

    def move(self, pos):
        """new local positions (ghost images follow, Verlet list kept)"""
        self.atoms.upload(POS, pos)
        self.ghosts.update_ghost_atoms(self.atoms, self.box)


def make_force_function():
    import torch

    class LJForce(torch.autograd.Function):
        """F = LJForce.apply(pos, eps_tab, sigma_tab, snapshot); backward runs the adjoint kernel"""

        @staticmethod
        def forward(ctx, pos, eps_tab, sig_tab, snap):
            snap.set_params(eps_tab.detach().cpu().numpy(), sig_tab.detach().cpu().numpy())
            snap.move(pos.detach().cpu().numpy())
            F = torch.from_numpy(snap.forces()).to(pos.device)
            ctx.snap = snap
            return F

        @staticmethod
        def backward(ctx, gF):
            dx, de, ds = ctx.snap.adjoint(gF.detach().cpu().numpy().astype(np.float64))
            # the kernel returns the derivative w.r.t. the shared pair parameter in both
            # symmetric entries; split it so the table-level chain rule (eps_ab = eps_ba
            # built from per-type values) counts it once
            nt = ctx.snap.num_types
            de = de.reshape(nt, nt).copy()
            ds = ds.reshape(nt, nt).copy()
            off = ~np.eye(nt, dtype=bool)
            de[off] *= 0.5
            ds[off] *= 0.5
            dev = gF.device
            return (torch.from_numpy(dx).to(dev), torch.from_numpy(de.reshape(-1)).to(dev),
                    torch.from_numpy(ds.reshape(-1)).to(dev), None)

    return LJForce


def equilibrated_snapshots(ctx, n_snap, cells, rho, frac_b, temperature, seed, steps_between=100,
                           eps_types=(1.0, 0.5), sigma_types=(1.0, 0.88), rc=2.5):
    """seeded two-type mixture relaxed with Langevin dynamics; returns position arrays,
    types and the box edge"""
    from .drivers import LJSystem
    from .lattice import fcc, maxwell_boltzmann, rng
    pos, L = fcc(cells, rho)
    n = pos.shape[0]
    types = (rng(seed).random(n) < frac_b).astype(np.int64)
    eA, eB = eps_types
    sA, sB = sigma_types
    eps = np.array([eA, np.sqrt(eA * eB), np.sqrt(eA * eB), eB])
    sig = np.array([sA, 0.5 * (sA + sB), 0.5 * (sA + sB), sB])
    s = LJSystem(ctx, pos, maxwell_boltzmann(n, temperature, seed), L, rc=rc, sigma=sig, epsilon=eps,
                 langevin=True, gamma=1.0, temperature=temperature, seed=seed, types=types, num_types=2,
                 type_stride=2)
    out = []
    s.run(200, 0.005)
    for _ in range(n_snap):
        s.run(steps_between, 0.005)
        p = s.atoms.download(POS, s.atoms.nlocal)
        t = s.atoms.download(3, s.atoms.nlocal)
        out.append((np.mod(p, L), t))
    return out, L
