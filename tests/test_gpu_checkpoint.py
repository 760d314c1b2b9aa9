"""Checkpoint and resume of single systems on the GPU (mrmd_amd.checkpoint).

* LJSystem (Langevin NVT, cell-list permutation on every rebuild): the run that goes on
  after a dump and a system restored from the dump take bit-identical steps (positions,
  velocities, forces, noise counter, rebuild count); restore-then-dump gives back the
  dumped particle datasets exactly;
* AdResSSystem: LJ_IdealGas run counter and compensation histograms, thermodynamic force,
  density profile and sample count come back exactly; the restored run follows the
  original to 1e-9 in positions (ghost-partner forces are added by atomics, so the order of
  additions can differ between any two runs).
"""
import numpy as np
import pytest

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


def test_lj_resume_bit_identical(M, ctx, tmp_path):
    from mrmd_amd import checkpoint
    from mrmd_amd.drivers import LJSystem
    pos, L = fcc(8, 0.8442)
    vel = maxwell_boltzmann(pos.shape[0], 1.2, 3)
    kw = dict(langevin=True, gamma=1.0, temperature=1.2, seed=99, permute=True)
    a = LJSystem(ctx, pos, vel, L, **kw)
    a.run(37, 0.005)
    path = str(tmp_path / "lj.npz")
    checkpoint.dump(path, a)
    a.run(41, 0.005)
    # a system built from other arrays: the checkpoint replaces its atoms and counters
    b = LJSystem(ctx, pos[:100], vel[:100], L, **kw)
    checkpoint.restore(path, b)
    assert b.step_count == 37 and b.langevin.call == 37 and b.atoms.nlocal == pos.shape[0]
    again = checkpoint.frame_of(b)
    saved = checkpoint.read_frame(path)
    for name, _ in checkpoint.DATASETS:
        k = f"particles/atoms/{name}/value"
        assert np.array_equal(again[k], saved[k]), k
    b.run(41, 0.005)
    assert a.rebuilds == b.rebuilds and a.rebuilds >= 4
    assert a.step_count == b.step_count == 78 and a.langevin.call == b.langevin.call
    for f in (M.POS, M.VEL, M.FORCE):
        assert np.array_equal(a.local(f), b.local(f)), f


def test_adress_resume(M, ctx, tmp_path):
    from mrmd_amd import checkpoint
    from mrmd_amd.drivers import AdResSSystem
    L = [40.0, 12.0, 12.0]
    g = np.random.default_rng(21)
    grid = np.stack(np.meshgrid(np.arange(32) * 1.25, np.arange(10) * 1.2, np.arange(10) * 1.2, indexing="ij"), -1)
    pos = np.mod(grid.reshape(-1, 3) + 0.3 + g.uniform(-0.05, 0.05, (3200, 3)), L)
    vel = maxwell_boltzmann(pos.shape[0], 1.0, 5)
    # the target density is a constructor parameter: both systems take the full system's
    kw = dict(cap=0.7, seed=11, tf_every=10, target_density=pos.shape[0] / float(np.prod(L)))

    def system(p, v):
        s = AdResSSystem(ctx, p, v, L, 8.0, 5.0, 1, **kw)
        s.ljig.set_intervals(3, 7)  # compensation sampling and updates inside the test's steps
        return s

    a = system(pos, vel)
    for _ in range(25):
        a.step(0.005)
    path = str(tmp_path / "adress.npz")
    checkpoint.dump(path, a)
    hist = [a.ljig.histogram(k) for k in range(3)]
    assert np.abs(hist[2]).max() > 0.0 and np.abs(a.tf.get_force()).max() > 0.0
    for _ in range(20):
        a.step(0.005)
    b = system(pos[:500], vel[:500])
    checkpoint.restore(path, b)
    assert b.ljig.run_counter == 25 and b.tf.samples == 5 and b.step_count == 25
    for k in range(3):
        assert np.array_equal(b.ljig.histogram(k), hist[k]), k
    saved = checkpoint.read_frame(path)
    assert np.array_equal(b.tf.get_force(), saved["thermodynamic_force/force"])
    assert np.array_equal(b.tf.get_density_profile(), saved["thermodynamic_force/density"])
    for _ in range(20):
        b.step(0.005)
    assert a.rebuilds == b.rebuilds and a.ljig.run_counter == b.ljig.run_counter == 45
    n = a.atoms.nlocal
    assert n == b.atoms.nlocal
    d = a.local(M.POS) - b.local(M.POS)
    d -= np.asarray(L) * np.round(d / np.asarray(L))
    assert np.abs(d).max() < 1e-9, np.abs(d).max()
    assert np.abs(a.tf.get_force() - b.tf.get_force()).max() < 1e-9
