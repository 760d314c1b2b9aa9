"""Ensemble engine on the GPU.

* a replica run inside the batch equals the same replica run alone, bit for bit, over 300
  steps of 10-step graphs (list rebuilds, thermodynamic-force updates every 50 steps, the
  compensation-energy sampling and update of step 0 included);
* graph replay equals plain launches bit for bit;
* forces of plain and AdResS replicas equal the oracle's ghost-layer algorithms (half list,
  atomic partner forces; LJ_IdealGas with molecules and multi-resolution ghosts) to 1e-12;
* a plain replica's trajectory follows the single-system GPU path (LJSystem);
* every replica reaches its thermostat temperature.
"""
import numpy as np
import pytest

import pyoracle as O
from mrmd_amd.lattice import maxwell_boltzmann, rng

pytestmark = pytest.mark.gpu

RC, SKIN, CAP = 2.5, 0.3, 0.7
SPECS = [(700, 0.8, 1.0, False), (1500, 0.5, 1.2, True), (600, 0.9, 0.8, False), (2500, 0.6, 1.0, True),
         (900, 0.3, 1.4, True), (3000, 0.85, 0.9, False)]


@pytest.fixture(scope="module")
def M():
    import mrmd_amd
    return mrmd_amd


@pytest.fixture(scope="module")
def ctx(M):
    c = M.Context(0)
    yield c
    c.close()


def _replicas():
    from mrmd_amd.ensemble import adress_box, cubic_box, replica
    reps, pos, vel = [], [], []
    g = rng(77)
    for k, (n, rho, t, ad) in enumerate(SPECS):
        box = adress_box(n, rho) if ad else cubic_box(n, rho)
        reps.append(replica(n, box, t, 100 + k, adress=ad, tf_every=50))
        pos.append(g.uniform(0.0, 1.0, (n, 3)) * np.asarray(box))
        vel.append(maxwell_boltzmann(n, t, 300 + k))
    return reps, pos, vel


def _ensemble(ctx, reps, pos, vel):
    from mrmd_amd.ensemble import Ensemble
    e = Ensemble(ctx, reps, RC, SKIN)
    e.set_lj(cap=CAP)
    e.upload(M_POS, pos)
    e.upload(M_VEL, vel)
    return e


M_POS, M_VEL = 0, 1


@pytest.fixture(scope="module")
def relaxed(ctx):
    """uniform-random starts relaxed 200 steps with the capped potential"""
    from mrmd_amd.ensemble import relax
    reps, pos, _ = _replicas()
    pos, vel = relax(ctx, reps, pos)
    return reps, pos, vel


def test_batch_equals_single_runs(ctx, relaxed):
    reps, pos, vel = relaxed
    e = _ensemble(ctx, reps, pos, vel)
    e.run(300, 10)
    P = e.positions()
    reb, steps = e.stats()
    assert steps == 300 and (reb > 1).all()
    for k in (1, 2, 3):
        one = _ensemble(ctx, [reps[k]], [pos[k]], [vel[k]])
        one.run(300, 10)
        assert np.array_equal(one.positions()[0], P[k]), k
        assert np.array_equal(one.velocities()[0], e.velocities()[k]), k
    # the thermodynamic force was updated (every 50 steps) in the AdResS replicas
    assert np.abs(e.tf(1, 0)).max() > 0.0


def test_graph_replay_equals_plain_launches(ctx, relaxed):
    reps, pos, vel = relaxed
    a = _ensemble(ctx, reps, pos, vel)
    b = _ensemble(ctx, reps, pos, vel)
    a.run(30, 10)
    b.run(30, 0)
    for x, y in zip(a.positions(), b.positions()):
        assert np.array_equal(x, y)


def _oracle_plain_forces(p, box):
    n = p.shape[0]
    a = O.Atoms.from_arrays(p)
    g = O.Ghost()
    rn = RC + SKIN
    g.create_xyz(a, O.make_box([0, 0, 0], box, rn))
    vl = O.Verlet()
    vl.build(a, a.nlocal + a.nghost, 0, a.nlocal, rn, 1.0, [-rn] * 3, [box[d] + rn for d in range(3)])
    lj = O.LJ([CAP], [RC], [1.0], [1.0], 1, shifted=True)
    lj.apply(a, vl)
    g.fold(a)
    return a.get(O.FORCE, n)


def _oracle_adress_forces(p, r):
    box = r["box"]
    rn = RC + SKIN
    n = p.shape[0]
    oa = O.Atoms.from_arrays(p)
    om = O.Molecules.for_each_atom(oa)
    obox = O.make_box([0, 0, 0], box, rn)
    slab = O.make_slab([0.5 * box[0], 0.5 * box[1], 0.5 * box[2]], r["atomistic_diameter"], r["hybrid_diameter"],
                       r["nu"])
    O.lib().or_update_molecules(om.h, oa.h, None)
    O.lib().or_multires_exchange_real(O.C.byref(obox), om.h, oa.h)
    og = O.Ghost()
    O.lib().or_multires_ghost_xyz(og.h, om.h, oa.h, O.C.byref(obox))
    O.lib().or_update_molecules(om.h, oa.h, O.C.byref(slab))
    nm = om.nlocal + om.nghost
    ov = O.Verlet()
    ov.build(om, nm, 0, om.nlocal, rn, 1.0, [-rn] * 3, [box[d] + rn for d in range(3)], molecules=True)
    olj = O.LJIdealGas([CAP], [RC], [1.0], [1.0], 1, shifted=True)
    na = oa.nlocal + oa.nghost
    oa.set(O.FORCE, np.zeros((na, 3)))
    O.lib().or_mol_set_force(om.h, 0.0)
    olj.run(om, ov, oa)
    O.lib().or_contribute_molecule_force(om.h, oa.h)
    og.fold(oa)
    return oa.get(O.FORCE, n)


def test_forces_match_oracle(ctx, relaxed):
    reps, pos, vel = relaxed
    e = _ensemble(ctx, reps, pos, vel)
    e.run(1, 0)  # one step: rebuild (wrap), forces at the stored positions, no TF / compensation yet
    P, F = e.positions(), e.forces()
    O.lib().or_set_threads(4)
    for k, r in enumerate(reps):
        ref = _oracle_adress_forces(P[k], r) if r["adress"] else _oracle_plain_forces(P[k], r["box"])
        scale = np.abs(ref).max()
        assert np.abs(F[k] - ref).max() <= 1e-12 * scale, (k, np.abs(F[k] - ref).max(), scale)


def test_plain_replica_follows_single_system_path(M, ctx, relaxed):
    from mrmd_amd.drivers import LJSystem
    reps, pos, vel = relaxed
    k = 0
    r = reps[k]
    one = _ensemble(ctx, [r], [pos[k]], [vel[k]])
    one.run(20, 0)
    s = LJSystem(ctx, pos[k], vel[k], r["box"], rc=RC, skin=SKIN, cap=CAP, shifted=True, langevin=True,
                 gamma=r["gamma"], temperature=r["temperature"], seed=r["seed"])
    s.run(20, r["dt"])
    ps = s.local(M.POS)
    L = np.asarray(r["box"])
    d = one.positions()[0] - ps
    d -= L * np.round(d / L)
    assert np.abs(d).max() < 1e-9


def test_thermostat_temperatures(ctx, relaxed):
    """after 4 time units (the relaxed structure releases its excess potential energy,
    which the thermostat removes with time constant 1/(2 gamma)) every replica sits at its
    temperature; averaged over the last 200 steps"""
    reps, pos, vel = relaxed
    e = _ensemble(ctx, reps, pos, vel)
    e.run(600, 10)
    ts = []
    for _ in range(10):
        e.run(20, 10)
        pe, ke = e.energies()
        assert np.isfinite(pe).all()
        ts.append(2.0 * ke / (3.0 * np.array([r["num_atoms"] for r in reps])))
    t = np.mean(ts, axis=0)
    for k, r in enumerate(reps):
        assert t[k] == pytest.approx(r["temperature"], rel=0.08), (k, t[k])


def test_global_fetch_equals_staged(ctx, relaxed):
    """tiles whose neighbourhood exceeds the LDS stage area read their records from global
    memory with the same arithmetic: capping the stage area at 64 records (every tile
    overflows) gives the same bits as the staged run over 30 steps"""
    reps, pos, vel = relaxed
    a = _ensemble(ctx, reps, pos, vel)
    b = _ensemble(ctx, reps, pos, vel)
    b.set_stage_cap(64)
    a.run(30, 10)
    b.run(30, 10)
    for x, y in zip(a.positions(), b.positions()):
        assert np.array_equal(x, y)


def test_lattice_start_runs(ctx):
    """a lattice start crowds some cells to ~3x the mean density (lattice planes aligned with
    cell faces), beyond the stage area sized from the mean: the run proceeds exactly"""
    from mrmd_amd.ensemble import Ensemble, cubic_box, replica
    n = 4000
    box = cubic_box(n, 0.8)
    sp = box[0] / np.ceil(n ** (1 / 3))
    m = int(np.ceil(n ** (1 / 3)))
    grid = np.stack(np.meshgrid(*[np.arange(m) * sp] * 3, indexing="ij"), -1).reshape(-1, 3)[:n]
    r = replica(n, box, 1.0, 3)
    e = Ensemble(ctx, [r], RC, SKIN)
    e.set_lj(cap=CAP)
    e.upload(M_POS, [grid + 1e-3])
    e.upload(M_VEL, [maxwell_boltzmann(n, 1.0, 4)])
    e.run(20, 10)
    f = _oracle_plain_forces(e.positions()[0], box)
    e2 = Ensemble(ctx, [r], RC, SKIN)
    e2.set_lj(cap=CAP)
    e2.upload(M_POS, e.positions())
    e2.upload(M_VEL, e.velocities())
    e2.run(1, 0)
    # one more step moved the atoms: compare the forces at the positions e2 used
    f2 = _oracle_plain_forces(e2.positions()[0], box)
    F = e2.forces()[0]
    assert np.abs(F - f2).max() <= 1e-12 * np.abs(f2).max()
    assert np.isfinite(f).all()


def test_checkpoint_resume_bit_identical(ctx, relaxed, tmp_path):
    """Ensemble.save / Ensemble.restore: the restored ensemble and the one that goes on
    after the save take bit-identical steps (graph replay, list rebuilds, the
    thermodynamic-force update of step 49 after the resume point)"""
    from mrmd_amd.ensemble import Ensemble
    reps, pos, vel = relaxed
    a = _ensemble(ctx, reps, pos, vel)
    a.run(45, 10)
    path = str(tmp_path / "ensemble.npz")
    a.save(path)
    a.run(35, 10)
    b = Ensemble.restore(ctx, path)
    assert b.stats()[1] == 45
    b.run(35, 10)
    ra, sa = a.stats()
    rb, sb = b.stats()
    assert sa == sb == 80 and np.array_equal(ra, rb)
    for x, y in zip(a.positions(), b.positions()):
        assert np.array_equal(x, y)
    for x, y in zip(a.velocities(), b.velocities()):
        assert np.array_equal(x, y)
    for k in (1, 3, 4):
        assert np.array_equal(a.tf(k, 0), b.tf(k, 0)) and np.array_equal(a.tf(k, 1), b.tf(k, 1)), k
    assert np.abs(a.tf(1, 0)).max() > 0.0
    pa, ka = a.energies()
    pb, kb = b.energies()
    assert np.array_equal(pa, pb) and np.array_equal(ka, kb)


def test_tiling_does_not_change_the_bits(ctx, relaxed):
    """two launch shapes (tile height target and workgroup size) give bit-identical runs: a
    row's entries follow its stencil columns and, in each, the z order of the records,
    whatever the tile that stages them"""
    from mrmd_amd.ensemble import Ensemble
    reps, pos, vel = relaxed
    out = []
    for target, block in ((1000, 128), (1800, 256)):
        e = Ensemble(ctx, reps, RC, SKIN)
        e.set_lj(cap=CAP)
        e.set_tiling(target, block)
        e.upload(0, pos)
        e.upload(1, vel)
        e.run(60, 10)
        out.append((e.positions(), e.velocities()))
    for k in range(len(reps)):
        np.testing.assert_array_equal(out[0][0][k], out[1][0][k])
        np.testing.assert_array_equal(out[0][1][k], out[1][1][k])
