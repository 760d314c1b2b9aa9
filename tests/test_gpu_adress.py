"""GPU parity of the AdResS path (slab weighting, molecule update, multi-resolution ghosts,
LJ_IdealGas, molecule force distribution, thermodynamic force) against the CPU oracle.

Index work (ghost order, corresponding-atom map, neighbour sets) is bit-exact.  Sums that
both sides accumulate with atomics in different orders (forces, energies, compensation
energies) agree to 1e-12 of their largest magnitude; counts are exact.
"""
import numpy as np
import pytest

import pyoracle as O
from mrmd_amd.lattice import uniform

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


def _close(a, b, rel=1e-12):
    scale = max(np.abs(b).max(), 1e-300)
    assert np.abs(a - b).max() <= rel * scale, (np.abs(a - b).max(), scale)


# LJ_IdealGas.test.cpp:34-220
@pytest.mark.parametrize("lam,scale", [(1.0, 1.0), (0.5, 0.5), (0.0, 0.0)])
def test_lj_ideal_gas_four_atoms(M, ctx, lam, scale):
    mols = M.Molecules(ctx, 2)
    mols.upload(M.M_POS, [[-0.5, 0, 0], [0.5, 0, 0]])
    mols.upload(M.M_OFFSET, [0, 2])
    mols.upload(M.M_NUMATOMS, [2, 2])
    mols.upload(M.M_MODLAMBDA, [lam, lam])
    mols.set_counts(2, 0)
    pos = np.array([[-0.5, -0.5, 0], [-0.5, 0.5, 0], [0.5, -0.5, 0], [0.5, 0.5, 0]])
    f0 = 2.0 if lam == 0.0 else 0.0
    atoms = M.Atoms.from_arrays(ctx, pos, relmass=np.full(4, 0.5), force=np.full((4, 3), f0))
    vl = M.HalfVerletList(ctx)
    vl.build(mols, 0, 2, 2.0, 1.0, [-1, -1, -1], [1, 1, 1], 4, molecules=True)
    lj = M.LJ_IdealGas(ctx, 0.0, 2.25, 0.9, 2.0, True)
    lj.run(mols, vl, atoms)
    F = atoms.download(M.FORCE)
    if lam == 0.0:
        assert np.all(F == 2.0)
        return
    fx, fy = 0.22156665 * scale, 1.3825009 * scale
    np.testing.assert_allclose(F, [[-fx, fy, 0], [-fx, -fy, 0], [fx, fy, 0], [fx, -fy, 0]], rtol=2e-7,
                               atol=1e-12)


def test_slab_and_histograms(M, ctx):
    slab = M.Slab([10, 0, 0], 4.0, 2.0, 1)
    os_ = O.make_slab([10, 0, 0], 4.0, 2.0, 1)
    xs = np.linspace(0, 20, 2001)
    g = slab(ctx, xs)
    for k, x in enumerate(xs):
        out = np.zeros(5)
        O.lib().or_slab_eval(O.C.byref(os_), x, O.ptr(out))
        np.testing.assert_allclose(g[k], out, rtol=1e-14, atol=1e-15)
    d = np.random.default_rng(0).random((23, 3))
    for periodic in (0, 1):
        og = np.zeros_like(d)
        O.lib().or_hist_gradient(0.0, 23.0, 23, 3, O.ptr(d), periodic, O.ptr(og))
        np.testing.assert_array_equal(M.hist_gradient(ctx, 0.0, 23.0, d, bool(periodic)), og)
        os2 = np.zeros_like(d)
        O.lib().or_hist_smoothen(0.0, 23.0, 23, 3, O.ptr(d), 2.0, 3.0, periodic, O.ptr(os2))
        np.testing.assert_allclose(M.hist_smoothen(ctx, 0.0, 23.0, d, 2.0, 3.0, bool(periodic)), os2, rtol=1e-14)


def _grid_fixture(M, ctx, apm=2):
    mp, ap = [], []
    for x in range(3):
        for y in range(3):
            for z in range(3):
                mp.append((x + 0.5, y + 0.5, z + 0.5))
                for i in range(apm):
                    ap.append((x + 0.5 + 0.1 * i, y + 0.5 + 0.2 * i, z + 0.5 + 0.3 * i))
    gm = M.Molecules(ctx, 27)
    gm.upload(M.M_POS, mp)
    gm.upload(M.M_OFFSET, np.arange(27) * apm)
    gm.upload(M.M_NUMATOMS, np.full(27, apm))
    gm.set_counts(27, 0)
    ga = M.Atoms.from_arrays(ctx, np.array(ap))
    om = O.Molecules(27)
    om.set(O.M_POS, mp)
    om.set(O.M_OFFSET, np.arange(27) * apm)
    om.set(O.M_NUMATOMS, np.full(27, apm))
    om.set_counts(27, 0)
    oa = O.Atoms.from_arrays(np.array(ap))
    return gm, ga, om, oa


def test_multires_ghosts_bit_exact(M, ctx):
    gm, ga, om, oa = _grid_fixture(M, ctx)
    gbox = M.make_box([0, 0, 0], [3, 3, 3], 0.7)
    obox = O.make_box([0, 0, 0], [3, 3, 3], 0.7)
    gl = M.GhostLayer(ctx)
    gl.multires_create_ghost_atoms(gm, ga, gbox)
    og = O.Ghost()
    O.lib().or_multires_ghost_xyz(og.h, om.h, oa.h, O.C.byref(obox))
    assert gm.nghost == 5 ** 3 - 27 == om.nghost
    assert ga.nghost == (5 ** 3 - 27) * 2 == oa.nghost
    nm, na = 27 + gm.nghost, 54 + ga.nghost
    np.testing.assert_array_equal(gm.download(M.M_POS, nm), om.get(O.M_POS, nm))
    np.testing.assert_array_equal(gm.download(M.M_OFFSET, nm), om.get(O.M_OFFSET, nm))
    np.testing.assert_array_equal(ga.download(M.POS, na), oa.get(O.POS, na))
    np.testing.assert_array_equal(gl.corresponding(na), og.corr(na))


@pytest.mark.parametrize("edges,ghost", [((9.0, 6.0, 11.0), 1.8), ((3.0, 7.0, 8.0), 1.8)])
def test_multires_three_axis_ghosts_equal_axis_by_axis(M, ctx, edges, ghost):
    """mrmd_multires_halo_create makes the three axes' ghost molecules (with their atoms) in
    one pass over the real molecules; molecules (positions, lambdas, atom offsets and counts),
    atoms (positions, velocities, masses) and correspondingRealAtom equal the axis-by-axis
    exchange's and the oracle's bit for bit: molecules of one to three atoms, a box thinner
    than two ghost widths along x, molecules exactly on the inner bounds"""
    rng = np.random.default_rng(5)
    L = np.array(edges)
    nm = 300
    mp = rng.uniform(0.0, 1.0, (nm, 3)) * L
    mp[0, 0], mp[1, 1], mp[2, 2] = ghost, L[1] - ghost, 0.0
    napm = rng.integers(1, 4, nm)
    off = np.concatenate([[0], np.cumsum(napm)[:-1]])
    ap = np.concatenate([mp[k] + rng.normal(0.0, 0.2, (napm[k], 3)) for k in range(nm)])
    vel = rng.normal(0.0, 1.0, ap.shape)
    mass = rng.uniform(0.5, 2.0, ap.shape[0])
    lam = rng.uniform(0.0, 1.0, nm)
    box = M.make_box([0, 0, 0], L, ghost)

    def make():
        gm = M.Molecules(ctx, nm)
        gm.upload(M.M_POS, mp)
        gm.upload(M.M_OFFSET, off)
        gm.upload(M.M_NUMATOMS, napm)
        gm.upload(M.M_LAMBDA, lam)
        gm.set_counts(nm, 0)
        return gm, M.Atoms.from_arrays(ctx, ap, vel=vel, mass=mass)

    ma, aa = make()
    mb, ab = make()
    ga, gb = M.GhostLayer(ctx), M.GhostLayer(ctx)
    ga.multires_create_ghost_atoms(ma, aa, box)
    for d in range(3):
        gb.multires_create_ghost_atoms_axis(mb, ab, box, d, reset=(d == 0))
    assert ma.nghost == mb.nghost > 0 and aa.nghost == ab.nghost
    tm, ta = nm + ma.nghost, ap.shape[0] + aa.nghost
    for f in (M.M_POS, M.M_LAMBDA, M.M_OFFSET, M.M_NUMATOMS):
        np.testing.assert_array_equal(ma.download(f, tm), mb.download(f, tm))
    for f in (M.POS, M.VEL, M.MASS):
        np.testing.assert_array_equal(aa.download(f, ta), ab.download(f, ta))
    np.testing.assert_array_equal(ga.corresponding(ta), gb.corresponding(ta))
    om = O.Molecules(nm)
    om.set(O.M_POS, mp)
    om.set(O.M_OFFSET, off)
    om.set(O.M_NUMATOMS, napm)
    om.set_counts(nm, 0)
    oa = O.Atoms.from_arrays(ap)
    og = O.Ghost()
    obox = O.make_box([0, 0, 0], L, ghost)
    O.lib().or_multires_ghost_xyz(og.h, om.h, oa.h, O.C.byref(obox))
    assert om.nghost == ma.nghost and oa.nghost == aa.nghost
    np.testing.assert_array_equal(ma.download(M.M_POS, tm), om.get(O.M_POS, tm))
    np.testing.assert_array_equal(ma.download(M.M_OFFSET, tm), om.get(O.M_OFFSET, tm))
    np.testing.assert_array_equal(aa.download(M.POS, ta), oa.get(O.POS, ta))
    np.testing.assert_array_equal(ga.corresponding(ta), og.corr(ta))


def _slab_system(M, ctx, n=6000, Lx=40.0, Lyz=12.0, seed=11):
    pos = uniform(n, [0, 0, 0], [Lx, Lyz, Lyz], seed)
    rng = np.random.default_rng(seed)
    types = (rng.random(n) < 0.3).astype(np.int64)
    return pos, types


@pytest.mark.parametrize("num_types,centre,queue", [(2, 20.0, None), (1, 20.0, None), (1, 3.0, None),
                                                   (1, 3.0, 0), (1, 3.0, 5)],
                         ids=["two-types-half-list", "one-type-tiled", "one-type-tiled-slab-across-boundary",
                              "one-type-tiled-no-queue", "one-type-tiled-queue-overflow"])
def test_adress_pipeline_matches_oracle(M, ctx, num_types, centre, queue):
    """one full AdResS force evaluation: multi-res ghosts, molecules, molecule list,
    LJ_IdealGas (sampling run), molecule -> atom forces, ghost fold, thermodynamic force.

    Two types run the half-list kernel; one type runs the tiled kernel.  With the slab
    centred at x = 3 the atomistic and hybrid regions cross the periodic boundary, so ghost
    images carry a lambda different from their real molecule's (ghosts at x < 0 are hybrid
    or atomistic, their reals near x = 40 coarse-grained): the tiled kernel must pair the
    same copies as the reference's half list.  Queue capacities 0 and 5 send the partner
    sides of ghost pairs through the rows' own second pass (all of them, or those past the
    first few)."""
    Lx, Lyz, rn = 40.0, 12.0, 2.8
    pos, types = _slab_system(M, ctx, Lx=Lx, Lyz=Lyz)
    if num_types == 1:
        types = np.zeros_like(types)
    n = pos.shape[0]
    gbox = M.make_box([0, 0, 0], [Lx, Lyz, Lyz], rn)
    obox = O.make_box([0, 0, 0], [Lx, Lyz, Lyz], rn)
    ga = M.Atoms.from_arrays(ctx, pos, type=types)
    oa = O.Atoms.from_arrays(pos, type=types)
    gm = M.Molecules.for_each_atom(ctx, ga)
    om = O.Molecules.for_each_atom(oa)
    slab = M.Slab([centre, 0, 0], 8.0, 5.0, 1)
    os_ = O.make_slab([centre, 0, 0], 8.0, 5.0, 1)
    gm.update(ga, None)
    O.lib().or_update_molecules(om.h, oa.h, None)
    gl = M.GhostLayer(ctx)
    og = O.Ghost()
    gl.multires_exchange_real_atoms(gm, ga, gbox)
    O.lib().or_multires_exchange_real(O.C.byref(obox), om.h, oa.h)
    gl.multires_create_ghost_atoms(gm, ga, gbox)
    O.lib().or_multires_ghost_xyz(og.h, om.h, oa.h, O.C.byref(obox))
    nm, na = gm.nlocal + gm.nghost, ga.nlocal + ga.nghost
    assert (nm, na) == (om.nlocal + om.nghost, oa.nlocal + oa.nghost)
    gm.update(ga, slab)
    O.lib().or_update_molecules(om.h, oa.h, O.C.byref(os_))
    for f in (M.M_POS,):
        np.testing.assert_array_equal(gm.download(f, nm), om.get(f, nm))
    _close(gm.download(M.M_MODLAMBDA, nm), om.get(O.M_MODLAMBDA, nm), 1e-14)
    vl = M.HalfVerletList(ctx)
    lo, hi = [-rn] * 3, [Lx + rn, Lyz + rn, Lyz + rn]
    vl.build(gm, 0, gm.nlocal, rn, 1.0, lo, hi, molecules=True)
    ov = O.Verlet()
    ov.build(om, nm, 0, om.nlocal, rn, 1.0, lo, hi, molecules=True)
    assert vl.num_pairs() == ov.total()
    if num_types == 2:
        cap, rc, sig, eps = [0.7] * 4, [2.5] * 4, [1.0, 0.95, 0.95, 0.9], [1.0, 0.8, 0.8, 0.6]
    else:
        cap, rc, sig, eps = [0.7], [2.5], [1.0], [1.0]
    glj = M.LJ_IdealGas(ctx, cap, rc, sig, eps, True, num_types=num_types)
    olj = O.LJIdealGas(cap, rc, sig, eps, num_types, shifted=True)
    if queue is not None:
        glj.set_queue(queue)
    ga.set_force(0.0)
    oa.set(O.FORCE, np.zeros((na, 3)))
    gm.set_force(0.0)
    O.lib().or_mol_set_force(om.h, 0.0)
    O.lib().or_set_threads(4)
    ctx.set_profiling(True)
    ctx.reset_stats()
    Eg = glj.run(gm, vl, ga)
    tiled_runs = ctx.kernel_stats("lj_idealgas_tiled")[0]
    ctx.set_profiling(False)
    assert tiled_runs == (1 if num_types == 1 else 0)
    Ec = olj.run(om, ov, oa)
    assert Eg == pytest.approx(Ec, rel=1e-12)
    if centre < 10.0:
        # the case is meant: some ghost images are non-CG while their reals are CG
        lam = gm.download(M.M_MODLAMBDA, nm)
        assert (lam[gm.nlocal:] > 0.0).any()
    _close(ga.download(M.FORCE, na), oa.get(O.FORCE, na))
    _close(gm.download(M.M_FORCE, nm), om.get(O.M_FORCE, nm))
    for which in (0, 1, 2):
        _close(glj.histogram(which), olj.hist(which), 1e-11)
    # molecule forces to atoms, ghost fold
    gm.contribute_force(ga)
    O.lib().or_contribute_molecule_force(om.h, oa.h)
    gl.contribute_back_ghost_to_real(ga)
    og.fold(oa)
    _close(ga.download(M.FORCE, ga.nlocal), oa.get(O.FORCE, oa.nlocal))
    # thermodynamic force: sample, update (symmetric, periodic), apply interpolated
    gtf = M.ThermodynamicForce(ctx, [0.5, 0.5], gbox, 1.0, [1.0, 1.0], True, True)
    otf = O.ThermodynamicForce([0.5, 0.5], obox, 1.0, [1.0, 1.0], sym=True, periodic=True)
    for _ in range(3):
        gtf.sample(ga)
        otf.sample(oa)
    np.testing.assert_array_equal(gtf.get_density_profile(), otf.density())
    hy = dict(kind=M.PRED_SLAB_HY, axis=0, a=Lx / 2, b=4.0, c=5.0)
    gtf.update(2.0, 3.0, M.pred1(**hy))
    otf.update(2.0, 3.0, O.pred1(**hy))
    _close(gtf.get_force(), otf.force(), 1e-13)
    gtf.apply_interpolated(ga)
    otf.apply_interpolated(oa)
    _close(ga.download(M.FORCE, ga.nlocal), oa.get(O.FORCE, oa.nlocal))
    gl_, gr_ = gtf.mu()
    ol_, or_ = otf.mu()
    np.testing.assert_allclose(gl_, ol_, rtol=1e-12, atol=1e-14)
    np.testing.assert_allclose(gr_, or_, rtol=1e-12, atol=1e-14)


def test_thermodynamic_force_reference_cases(M, ctx):
    # ThermodynamicForce.test.cpp:96-245 on the device
    box = M.make_box([0, 0, 0], [10, 1, 1], 1.0)
    tf = M.ThermodynamicForce(ctx, [1.0], box, 1.0, [1.0])
    tf.set_force(np.arange(10, dtype=float).reshape(-1, 1))
    pts = np.array([[i + 0.5, 0, 0] for i in range(10) for _ in range(i)])
    a = M.Atoms.from_arrays(ctx, pts)
    tf.sample(a)
    np.testing.assert_array_equal(tf.get_density_profile()[:, 0], np.arange(10.0))
    tf.update(1.0, 0.0)
    np.testing.assert_allclose(tf.get_force()[:, 0], np.arange(10.0) - 1.0)
    with pytest.raises(M.MrmdError):
        tf.update(1.0, 0.0)
    u = M.Atoms.from_arrays(ctx, np.array([[i / 10.0, 0, 0] for i in range(100)]))
    tf.set_force(np.arange(10, dtype=float).reshape(-1, 1))
    tf.apply_interpolated(u, M.pred1(M.PRED_OPEN_RANGE, 0, 0.5, 4.5))
    x = u.download(M.POS)[:, 0]
    inside = (x > 0.5) & (x < 4.5)
    np.testing.assert_allclose(u.download(M.FORCE)[:, 0], np.where(inside, x - 0.5, 0.0), atol=1e-12)


def test_thermo_force_dump_restore(M, ctx, tmp_path):
    """io/ThermoForce.test.cpp:25-120 through the Python mirror: forces (bin + 1)(type + 1)
    on 100 bins of subdomain (1, 2, 3)-(4, 6, 8), one type and five types"""
    from mrmd_amd.io import dump_thermo_force, restore_thermo_force
    box = M.make_box([1, 2, 3], [4, 6, 8], 0.5)
    for nt in (1, 5):
        tf1 = M.ThermodynamicForce(ctx, [1.0] * nt, box, 3.0 / 100, [1.0] * nt)
        assert tf1.num_bins == 100
        f = np.outer(np.arange(1, 101), np.arange(1, nt + 1)).astype(float)
        tf1.set_force(f)
        path = str(tmp_path / f"force{nt}.txt")
        if nt == 1:
            dump_thermo_force(path, tf1, 0)
        else:
            dump_thermo_force(path, tf1)
        tf2 = restore_thermo_force(ctx, path, box, [1.0] * nt, [1.0] * nt)
        np.testing.assert_allclose(tf2.grid(), tf1.grid(), rtol=1e-6)
        np.testing.assert_array_equal(tf2.get_force(), f)


def test_spherical_matches_oracle(M, ctx):
    """weighting_function::Spherical (Spherical.hpp:37-76) on the device against the oracle
    on points spread over the atomistic core, the hybrid shell and beyond (exponents 1-3)"""
    g = np.random.default_rng(9)
    c = np.array([2.0, 3.0, 4.0])
    pts = c + g.normal(size=(3000, 3)) * 2.5
    for n in (1, 2, 3):
        out = M.Spherical(c, 2.0, 2.0, n)(ctx, pts)
        ref = np.zeros_like(out)
        for k, p in enumerate(pts):
            O.lib().or_spherical_eval(O.ptr(c), 2.0, 2.0, n, p[0], p[1], p[2], O.ptr(ref[k]))
        np.testing.assert_allclose(out, ref, rtol=1e-13, atol=1e-15)
        assert (out[:, 0] == 1.0).any() and (out[:, 0] == 0.0).any() and ((out[:, 0] > 0) & (out[:, 0] < 1)).any()


def test_adress_trajectory_matches_oracle(M, ctx):
    """40 steps of the single-system AdResS driver (drivers.AdResSSystem) against the oracle's
    step loop (pyoracle.AdResSRun) from the same start: Langevin noise (the same Philox
    stream), rebuild steps, multi-resolution exchange and ghosts, the tiled LJ_IdealGas,
    molecule forces, ghost fold, thermodynamic force sampled every step and updated every 10
    steps over the hybrid region.  Positions agree to 1e-9 after 40 steps (the pair sums
    differ in order, so trajectories separate slowly from the ~1e-15 level)."""
    from mrmd_amd.drivers import AdResSSystem
    from mrmd_amd.lattice import maxwell_boltzmann
    L = [40.0, 12.0, 12.0]
    g = np.random.default_rng(21)
    grid = np.stack(np.meshgrid(np.arange(32) * 1.25, np.arange(10) * 1.2, np.arange(10) * 1.2, indexing="ij"), -1)
    pos = np.mod(grid.reshape(-1, 3) + 0.3 + g.uniform(-0.05, 0.05, (3200, 3)), L)
    vel = maxwell_boltzmann(pos.shape[0], 1.0, 5)
    kw = dict(cap=0.7, seed=11, tf_every=10)
    gsys = AdResSSystem(ctx, pos, vel, L, 8.0, 5.0, 1, **kw)
    osys = O.AdResSRun(pos, vel, L, 8.0, 5.0, 1, **kw)
    for _ in range(40):
        gsys.step(0.005)
        osys.step(0.005)
    assert gsys.rebuilds == osys.rebuilds and gsys.rebuilds >= 2
    n = gsys.atoms.nlocal
    assert n == osys.atoms.nlocal
    d = gsys.local(M.POS) - osys.atoms.get(O.POS, n)
    d -= np.asarray(L) * np.round(d / np.asarray(L))
    assert np.abs(d).max() < 1e-9, np.abs(d).max()
    dv = gsys.local(M.VEL) - osys.atoms.get(O.VEL, n)
    assert np.abs(dv).max() < 1e-7, np.abs(dv).max()
    np.testing.assert_allclose(gsys.tf.get_force(), osys.tf.force(), rtol=1e-9, atol=1e-10)
    assert np.abs(osys.tf.force()).max() > 0.0


def test_adress_native_loop_equals_step_calls(M, ctx):
    """mrmd_adress_md_run (the AdResS step in C++) against the same calls made step by step
    from Python: same rebuilds, noise counter, run counter and thermodynamic-force updates;
    positions agree to 1e-9 (ghost-partner forces are added by atomics, so the order of
    additions, and hence the last bits, can differ between any two runs)"""
    from mrmd_amd.drivers import AdResSSystem
    from mrmd_amd.lattice import maxwell_boltzmann
    L = [40.0, 12.0, 12.0]
    g = np.random.default_rng(21)
    grid = np.stack(np.meshgrid(np.arange(32) * 1.25, np.arange(10) * 1.2, np.arange(10) * 1.2, indexing="ij"), -1)
    pos = np.mod(grid.reshape(-1, 3) + 0.3 + g.uniform(-0.05, 0.05, (3200, 3)), L)
    vel = maxwell_boltzmann(pos.shape[0], 1.0, 5)
    kw = dict(cap=0.7, seed=11, tf_every=10)
    a = AdResSSystem(ctx, pos, vel, L, 8.0, 5.0, 1, **kw)
    b = AdResSSystem(ctx, pos, vel, L, 8.0, 5.0, 1, **kw)
    for _ in range(40):
        a.step(0.005)
    b.run(15, 0.005)
    b.run(25, 0.005)
    assert a.rebuilds == b.rebuilds and a.rebuilds >= 2
    assert a.step_count == b.step_count == 40 and a.langevin.call == b.langevin.call == 40
    assert a.ljig.run_counter == b.ljig.run_counter == 40 and a.tf.samples == b.tf.samples
    assert a.energy == pytest.approx(b.energy, rel=1e-10)
    d = a.local(M.POS) - b.local(M.POS)
    d -= np.asarray(L) * np.round(d / np.asarray(L))
    assert np.abs(d).max() < 1e-9, np.abs(d).max()
    assert np.abs(a.tf.get_force() - b.tf.get_force()).max() < 1e-9


def test_adress_loop_builds_rows_near_the_weighted_region_only(M, ctx):
    """The native AdResS loop builds list rows only in tiles within reach of the slab's
    atomistic and hybrid regions (x band widened by half the skin): on a long slab most
    tiles are coarse-grained and skipped.  Its trajectory equals the step-by-step calls,
    whose lists are full, and the banded list refuses uses other than LJ_IdealGas."""
    from mrmd_amd.drivers import AdResSSystem
    from mrmd_amd.lattice import maxwell_boltzmann
    L = [80.0, 12.0, 12.0]
    g = np.random.default_rng(5)
    grid = np.stack(np.meshgrid(np.arange(64) * 1.25, np.arange(10) * 1.2, np.arange(10) * 1.2, indexing="ij"), -1)
    pos = np.mod(grid.reshape(-1, 3) + 0.3 + g.uniform(-0.05, 0.05, (6400, 3)), L)
    vel = maxwell_boltzmann(pos.shape[0], 1.0, 9)
    kw = dict(cap=0.7, seed=13, tf_every=10)
    a = AdResSSystem(ctx, pos, vel, L, 8.0, 5.0, 1, **kw)
    b = AdResSSystem(ctx, pos, vel, L, 8.0, 5.0, 1, **kw)
    for _ in range(30):
        a.step(0.005)
    b.run(30, 0.005)
    assert a.rebuilds == b.rebuilds and a.rebuilds >= 2
    d = a.local(M.POS) - b.local(M.POS)
    d -= np.asarray(L) * np.round(d / np.asarray(L))
    assert np.abs(d).max() < 1e-9, np.abs(d).max()
    assert a.energy == pytest.approx(b.energy, rel=1e-10)
    with pytest.raises(M.MrmdError, match="weighted region"):
        b.vl.info()
