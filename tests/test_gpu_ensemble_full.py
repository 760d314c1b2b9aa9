"""The 256-replica ensemble at its full size: atom counts log-uniform in [500, 50000]
(2.72 M atoms), densities uniform in [0.3, 0.9], temperatures uniform in [0.7, 1.5], half
AdResS slabs (thermodynamic force updated every 100 steps), rc 2.5, skin 0.3, dt 0.005,
capped potential (r_cap 0.7), uniform-random starts relaxed 200 steps, then 2,000 steps in
10-step graph blocks.

* Replicas of the ensemble equal the same replicas run alone, bit for bit, over the 2,000
  steps: the smallest and the largest plain replica and the smallest and the largest AdResS
  slab, each with its own rebuilds, thermodynamic-force updates and compensation sampling.
* Every replica sits at its own temperature over the last 500 steps (snapshots every 50
  steps; within 2 % plus three standard deviations of a single snapshot's kinetic
  temperature, sqrt(2 / (3 N)), which allows for the small replicas' fluctuations), and
  every replica rebuilt its list.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

STEPS = 2000


@pytest.fixture(scope="module")
def ctx():
    import mrmd_amd
    c = mrmd_amd.Context(0)
    yield c
    c.close()


@pytest.fixture(scope="module")
def run(ctx):
    from mrmd_amd.ensemble import Ensemble, ensemble_config, relax
    reps, pos, _ = ensemble_config(n_rep=256, seed=2024)
    pos, vel = relax(ctx, reps, pos)
    e = Ensemble(ctx, reps)
    e.set_lj(cap=0.7)
    e.upload(0, pos)
    e.upload(1, vel)
    e.run(STEPS - 500, 10)
    temps = []
    for _ in range(10):
        e.run(50, 10)
        v = e.velocities()
        temps.append([np.mean(np.sum(x * x, axis=1)) / 3.0 for x in v])  # m = 1: 2 Ek / (3 N)
    return reps, pos, vel, e, np.array(temps)


def test_replicas_equal_their_single_runs(ctx, run):
    from mrmd_amd.ensemble import Ensemble
    reps, pos, vel, e, _ = run
    P, V = e.positions(), e.velocities()
    n = np.array([r["num_atoms"] for r in reps])
    ad = np.array([r["adress"] for r in reps], bool)
    picks = {int(np.argmin(np.where(ad, np.inf, n))), int(np.argmax(np.where(ad, -1, n))),
             int(np.argmin(np.where(ad, n, np.inf))), int(np.argmax(np.where(ad, n, -1)))}
    assert len(picks) == 4
    for k in sorted(picks):
        one = Ensemble(ctx, [reps[k]])
        one.set_lj(cap=0.7)
        one.upload(0, [pos[k]])
        one.upload(1, [vel[k]])
        one.run(STEPS, 10)
        assert np.array_equal(one.positions()[0], P[k]), (k, reps[k]["num_atoms"], reps[k]["adress"])
        assert np.array_equal(one.velocities()[0], V[k]), k


def test_every_replica_at_its_temperature(run):
    reps, _, _, e, temps = run
    target = np.array([r["temperature"] for r in reps])
    n = np.array([r["num_atoms"] for r in reps])
    mean = temps.mean(axis=0)
    rel = np.abs(mean - target) / target
    tol = 0.02 + 3.0 * np.sqrt(2.0 / (3.0 * n))
    bad = np.nonzero(rel > tol)[0]
    assert bad.size == 0, [(int(k), int(n[k]), mean[k], target[k]) for k in bad]
    reb, steps = e.stats()
    assert steps == STEPS and (reb >= 2).all()
