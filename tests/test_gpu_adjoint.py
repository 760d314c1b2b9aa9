"""Adjoint LJ kernel on the GPU: against the CPU adjoint (oracle), against FP64 central
finite differences (h = 1e-6, agreement 1e-6), its cost relative to the forward pass
(<= 3x), and a force-matching fit through the torch.autograd.Function harness.

Workload of the force-matching configuration: two-type LJ mixture, N = 32,000, 80/20 A/B,
density 0.8, T = 1.0, Lorentz-Berthelot mixing, eps_AA = 1.0, eps_BB = 0.5, sigma_AA = 1.0,
sigma_BB = 0.88, rc = 2.5 (shifted); snapshots from seeded Langevin runs; upstream dL/dF
drawn i.i.d. N(0, 1).
"""
import numpy as np
import pytest

import pyoracle as O
from mrmd_amd.lattice import fcc

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def M():
    import mrmd_amd
    return mrmd_amd


@pytest.fixture(scope="module")
def ctx(M):
    c = M.Context(0)
    yield c
    c.close()


def _lb(eA, eB, sA, sB):
    eAB = np.sqrt(eA * eB)
    sAB = 0.5 * (sA + sB)
    return np.array([eA, eAB, eAB, eB]), np.array([sA, sAB, sAB, sB])


def _mixture(cells=10, rho=0.8, seed=5):
    pos, L = fcc(cells, rho)
    rng = np.random.default_rng(seed)
    pos = np.mod(pos + rng.uniform(-0.12, 0.12, pos.shape), L)
    types = (rng.random(pos.shape[0]) < 0.2).astype(np.int64)
    return pos, types, L


def test_adjoint_matches_cpu_reference_gradient(M, ctx):
    from mrmd_amd.adjoint import Snapshot
    pos, types, L = _mixture()
    n = pos.shape[0]
    eps, sig = _lb(1.0, 0.5, 1.0, 0.88)
    snap = Snapshot(ctx, pos, types, L, 2, eps, sig)
    up = np.random.default_rng(2).standard_normal((n, 3))
    Fg = snap.forces()
    dx, de, ds = snap.adjoint(up)
    # CPU reference
    O.lib().or_set_threads(4)
    a = O.Atoms.from_arrays(pos, type=types)
    g = O.Ghost()
    g.create_xyz(a, O.make_box([0, 0, 0], [L] * 3, 2.8))
    vl = O.Verlet()
    vl.build(a, a.nlocal + a.nghost, 0, a.nlocal, 2.8, 1.0, [-2.8] * 3, [L + 2.8] * 3)
    lj = O.LJ([0.0] * 4, [2.5] * 4, sig, eps, 2, shifted=True, type_stride=2)
    lj.apply(a, vl)
    g.fold(a)
    Fc = a.get(O.FORCE, n)
    assert np.abs(Fg - Fc).max() <= 1e-12 * np.abs(Fc).max()
    cdx, cde, cds = np.zeros((n, 3)), np.zeros(4), np.zeros(4)
    O.lib().or_lj_adjoint(lj.h, a.h, vl.h, g.h, O.ptr(up), O.ptr(cdx), O.ptr(cde), O.ptr(cds))
    assert np.abs(dx - cdx).max() <= 1e-11 * np.abs(cdx).max()
    np.testing.assert_allclose(de, cde, rtol=1e-10)
    np.testing.assert_allclose(ds, cds, rtol=1e-10)


def test_adjoint_matches_finite_differences(M, ctx):
    from mrmd_amd.adjoint import Snapshot
    pos, types, L = _mixture(cells=6)
    n = pos.shape[0]
    eps, sig = _lb(1.0, 0.5, 1.0, 0.88)
    snap = Snapshot(ctx, pos, types, L, 2, eps, sig)
    up = np.random.default_rng(4).standard_normal((n, 3))
    snap.forces()
    dx, de, ds = snap.adjoint(up)
    h = 1e-6
    rng = np.random.default_rng(8)
    for k in rng.choice(n, 10, replace=False):
        for d in range(3):
            pp, pm = pos.copy(), pos.copy()
            pp[k, d] += h
            pm[k, d] -= h
            snap.move(pp)
            Fp = snap.forces()
            snap.move(pm)
            Fm = snap.forces()
            fd = np.sum(up * (Fp - Fm)) / (2 * h)
            assert dx[k, d] == pytest.approx(fd, rel=1e-6, abs=1e-7)
    snap.move(pos)
    for which, grad in (("eps", de), ("sig", ds)):
        for t, entries in ((0, [0]), (1, [1, 2]), (3, [3])):
            ep, em, sp, sm = eps.copy(), eps.copy(), sig.copy(), sig.copy()
            for e in entries:
                if which == "eps":
                    ep[e] += h
                    em[e] -= h
                else:
                    sp[e] += h
                    sm[e] -= h
            snap.set_params(ep, sp)
            Fp = snap.forces()
            snap.set_params(em, sm)
            Fm = snap.forces()
            fd = np.sum(up * (Fp - Fm)) / (2 * h)
            assert grad[t] == pytest.approx(fd, rel=1e-6), (which, t)
    snap.set_params(eps, sig)


def test_adjoint_cost_and_force_matching_fit(M, ctx):
    """64 snapshots of the N = 32,000 mixture: adjoint kernel time <= 3x the forward
    kernel (and the whole adjoint call, host arrays in and out, <= 3x the whole force call);
    Adam recovers parameters perturbed by +10% from forces alone, cycling over the 64
    snapshots two per step"""
    import time

    import torch

    from mrmd_amd.adjoint import Snapshot, equilibrated_snapshots, lorentz_berthelot, make_force_function
    nsnap = 64
    snaps_raw, L = equilibrated_snapshots(ctx, nsnap, 20, 0.8, 0.2, 1.0, 77, steps_between=50)
    eps0, sig0 = _lb(1.0, 0.5, 1.0, 0.88)
    snaps = [Snapshot(ctx, p, t, L, 2, eps0, sig0) for p, t in snaps_raw]
    assert snaps[0].n == 32000
    # cost ratio from HIP events on the library stream
    up = np.random.default_rng(1).standard_normal((32000, 3))
    ctx.set_profiling(True)
    ctx.reset_stats()
    for s in snaps:
        for _ in range(3):
            s.forces()
            s.adjoint(up)
    nf, tf = ctx.kernel_stats("lj_force")
    na, ta = ctx.kernel_stats("lj_adjoint")
    ctx.set_profiling(False)
    ratio = (ta / na) / (tf / nf)
    print(f"adjoint/forward kernel time ratio {ratio:.3f} ({ta / na:.4f} ms vs {tf / nf:.4f} ms)")
    assert ratio <= 3.0
    ctx.synchronize()
    t0 = time.perf_counter()
    for s in snaps:
        s.forces()
    t_fwd = time.perf_counter() - t0
    t0 = time.perf_counter()
    for s in snaps:
        s.adjoint(up)
    t_adj = time.perf_counter() - t0
    print(f"whole calls: adjoint {t_adj / nsnap * 1e3:.3f} ms, forward {t_fwd / nsnap * 1e3:.3f} ms")
    assert t_adj <= 3.0 * t_fwd
    # targets from +10 % parameters
    target = _lb(1.1, 0.55, 1.1, 0.968)
    F_target = []
    for s in snaps:
        s.set_params(*target)
        F_target.append(torch.from_numpy(s.forces()))
    LJForce = make_force_function()
    theta = torch.tensor(np.log([1.0, 0.5, 1.0, 0.88]), dtype=torch.float64, requires_grad=True)
    opt = torch.optim.Adam([theta], lr=0.01)
    pos_t = [torch.from_numpy(p) for p, _ in snaps_raw]
    first = None
    for it in range(200):
        opt.zero_grad()
        p = torch.exp(theta)
        e_tab, s_tab = lorentz_berthelot(p[:2], p[2:])
        loss = 0.0
        for k in (it % nsnap, (it + 37) % nsnap):
            F = LJForce.apply(pos_t[k], e_tab, s_tab, snaps[k])
            loss = loss + torch.mean((F - F_target[k]) ** 2)
        loss.backward()
        opt.step()
        first = first if first is not None else loss.item()
    final = torch.exp(theta).detach().numpy()
    print("fitted", final, "loss", first, "->", loss.item())
    np.testing.assert_allclose(final, [1.1, 0.55, 1.1, 0.968], rtol=0.02)
    assert loss.item() < 1e-3 * first
