"""GPU time-loop parity: the plain LJ NVE configuration (N = 4,000 FCC, rho = 0.8442,
T = 0.72, rc = 2.5 shifted, skin 0.3, dt = 0.005, 1,000 velocity-Verlet steps with the
reference's rebuild rule) against the CPU oracle's restatement of the same loop.

Per-pair forces agree to ~1e-16 relative, so trajectories separate only through the
chaos of the liquid: the early energy series agree to 1e-9, the rebuild count exactly,
and the NVE drift (slope of E_total per atom) to within 2e-6 per tau.
"""
import numpy as np
import pytest

import pyoracle as O
from mrmd_amd.lattice import fcc, maxwell_boltzmann

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


def _cfg1():
    pos, L = fcc(10, 0.8442)
    vel = maxwell_boltzmann(pos.shape[0], 0.72, 2024)
    return pos, vel, L


def _oracle_run(pos, vel, L, nsteps, every, langevin=0, permute=0):
    a = O.Atoms.from_arrays(pos, vel=vel)
    box = O.make_box([0, 0, 0], [L] * 3, 2.8)
    nrec = (nsteps + every - 1) // every
    out = np.zeros((nrec, 3))
    prm = O.md_params(nsteps=nsteps, energy_every=every, langevin=langevin, permute=permute, shifted=1,
                      gamma=1.0, temperature=1.0, seed=1234)
    O.lib().or_set_threads(4)
    reb = O.lib().or_run_md(a.h, O.C.byref(box), O.C.byref(prm), O.ptr(out))
    return out, reb, a


def test_nve_config1_energy_series_and_drift(M, ctx):
    from mrmd_amd.drivers import LJSystem
    pos, vel, L = _cfg1()
    n = pos.shape[0]
    sysg = LJSystem(ctx, pos, vel, L)
    eg = sysg.run(1000, 0.005, energy_every=10)
    ec, reb_c, _ = _oracle_run(pos, vel, L, 1000, 10)
    assert sysg.rebuilds == reb_c
    tot_g = (eg[:, 0] + eg[:, 1]) / n
    tot_c = (ec[:, 0] + ec[:, 1]) / n
    # early agreement (before chaos amplifies the 1e-16 force differences)
    np.testing.assert_allclose(eg[:20, 0], ec[:20, 0], rtol=1e-9)
    np.testing.assert_allclose(eg[:20, 1], ec[:20, 1], rtol=1e-9)
    # energy conservation and matching drift
    t = np.arange(len(tot_g)) * 10 * 0.005
    slope_g = np.polyfit(t, tot_g, 1)[0]
    slope_c = np.polyfit(t, tot_c, 1)[0]
    assert abs(tot_g - tot_g[0]).max() < 5e-3
    assert abs(slope_g - slope_c) < 2e-6 + 0.5 * abs(slope_c)


def test_langevin_permute_loop_matches_early_steps(M, ctx):
    from mrmd_amd.drivers import LJSystem
    pos, vel, L = fcc(6, 0.8442)[0], None, fcc(6, 0.8442)[1]
    vel = maxwell_boltzmann(pos.shape[0], 1.0, 7)
    sysg = LJSystem(ctx, pos, vel, L, langevin=True, gamma=1.0, temperature=1.0, seed=1234, permute=True)
    eg = sysg.run(30, 0.005, energy_every=1)
    ec, reb_c, a = _oracle_run(pos, vel, L, 30, 1, langevin=1, permute=1)
    assert sysg.rebuilds == reb_c
    np.testing.assert_allclose(eg[:, 0], ec[:, 0], rtol=1e-10)
    np.testing.assert_allclose(eg[:, 1], ec[:, 1], rtol=1e-10)
    # same permutation and state on both sides
    np.testing.assert_allclose(sysg.local(M.POS), a.get(O.POS, a.nlocal), rtol=0, atol=1e-10)


@pytest.mark.parametrize("langevin,permute,masses", [(False, False, False), (True, True, False), (True, True, True),
                                                    (False, True, True)],
                         ids=["nve", "nvt-permute", "nvt-permute-masses", "nve-permute-masses"])
def test_native_loop_equals_step_calls(M, ctx, langevin, permute, masses):
    """mrmd_md_run (the example loop in C++: fused post/pre passes, slot-order positions
    written by the integrator and the ghost update instead of a gather) against the same
    calls made step by step from Python: bit-identical positions, velocities and forces,
    same rebuilds and noise counter, across two native calls that continue one run"""
    from mrmd_amd.drivers import LJSystem
    pos, L = fcc(8, 0.8442)
    vel = maxwell_boltzmann(pos.shape[0], 1.2, 17)
    kw = dict(langevin=langevin, gamma=1.0, temperature=1.2, seed=5, permute=permute)
    a = LJSystem(ctx, pos, vel, L, **kw)
    b = LJSystem(ctx, pos, vel, L, **kw)
    if masses:
        m = 0.8 + 0.4 * np.random.default_rng(3).random(pos.shape[0])
        a.atoms.upload(M.MASS, m)
        b.atoms.upload(M.MASS, m)
    for _ in range(60):
        a.step(0.005)
    b.run(25, 0.005)
    b.run(35, 0.005)
    assert a.rebuilds == b.rebuilds and a.rebuilds >= 3 and a.step_count == b.step_count == 60
    assert a.max_disp == b.max_disp
    if langevin:
        assert a.langevin.call == b.langevin.call == 60
    assert a.atoms.nghost == b.atoms.nghost
    for f in (M.POS, M.VEL, M.FORCE):
        assert np.array_equal(a.local(f), b.local(f)), f
    # the ghost records' positions too (the native loop keeps positions in the list's slot
    # order between rebuilds and copies them back when the call returns)
    n = a.atoms.nlocal + a.atoms.nghost
    assert np.array_equal(a.atoms.download(M.POS, n), b.atoms.download(M.POS, n))
    # and their forces (zero on both sides; the native loop fills them once per ghost creation)
    assert np.array_equal(a.atoms.download(M.FORCE, n), b.atoms.download(M.FORCE, n))
