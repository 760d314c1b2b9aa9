"""GPU parity of the plain LJ hot path against the CPU oracle (calls go through the C ABI).

Tolerances: integer/index work (ghost order, cell assignment, neighbour sets) is bit-exact;
element-wise integrators and the Philox noise are bit-exact given identical inputs; pair
forces are compared with max_i |F_gpu - F_cpu| <= 1e-12 * max_i |F_cpu| (the device sums
both directions of every pair in another order than the half-list CPU loop).
"""
import numpy as np
import pytest

import pyoracle as O
from mrmd_amd.lattice import fcc, maxwell_boltzmann

pytestmark = pytest.mark.gpu

L_ESPP = 33.8585
F_TOL = 1e-12


@pytest.fixture(scope="module")
def M():
    import mrmd_amd
    return mrmd_amd


@pytest.fixture(scope="module")
def ctx(M):
    c = M.Context(0)
    yield c
    c.close()


def _both(M, ctx, pos, **kw):
    g = M.Atoms.from_arrays(ctx, pos, **kw)
    c = O.Atoms.from_arrays(pos, **{k: v for k, v in kw.items() if k != "charge"})
    return g, c


def _espp(M, ctx, espp_positions):
    gbox = M.make_box([0, 0, 0], [L_ESPP] * 3, 2.8)
    cbox = O.make_box([0, 0, 0], [L_ESPP] * 3, 2.8)
    g, c = _both(M, ctx, espp_positions)
    return gbox, cbox, g, c


def test_espp_ghosts_bit_exact(M, ctx, espp_positions):
    gbox, cbox, g, c = _espp(M, ctx, espp_positions)
    gl = M.GhostLayer(ctx)
    assert gl.create_ghost_atoms(g, gbox) == 22104
    og = O.Ghost()
    og.create_xyz(c, cbox)
    n = g.nlocal + g.nghost
    assert n == c.nlocal + c.nghost
    np.testing.assert_array_equal(g.download(M.POS, n), c.get(O.POS, n))
    np.testing.assert_array_equal(gl.corresponding(n), og.corr(n))


@pytest.mark.parametrize("edges,ghost", [((10.0, 7.0, 12.0), 2.0), ((3.0, 8.0, 9.0), 2.0), ((4.0, 3.5, 3.0), 2.2)])
def test_three_axis_ghosts_equal_axis_by_axis(M, ctx, edges, ghost):
    """mrmd_halo_create makes the three axes' ghosts in one pass over the real atoms; its
    records (positions, velocities, masses) and correspondingRealAtom equal those of the
    axis-by-axis exchange (mrmd_halo_create_axis x, y, z) and the oracle's bit for bit,
    including boxes thinner than two ghost widths (atoms copied both +L and -L along an
    axis) and atoms exactly on the inner bounds"""
    rng = np.random.default_rng(7)
    L = np.array(edges)
    n = 600
    pos = rng.uniform(0.0, 1.0, (n, 3)) * L
    pos[0, 0], pos[1, 1], pos[2, 2] = ghost, L[1] - ghost, 0.0
    pos[3] = [L[0] - ghost, ghost, L[2] - ghost]
    vel = rng.normal(0.0, 1.0, (n, 3))
    mass = rng.uniform(0.5, 2.0, n)
    box = M.make_box([0, 0, 0], L, ghost)
    a = M.Atoms.from_arrays(ctx, pos, vel=vel, mass=mass)
    b = M.Atoms.from_arrays(ctx, pos, vel=vel, mass=mass)
    ga, gb = M.GhostLayer(ctx), M.GhostLayer(ctx)
    na = ga.create_ghost_atoms(a, box)
    for d in range(3):
        gb.create_ghost_atoms_axis(b, box, d, reset=(d == 0))
    assert na == b.nghost > 0
    tot = n + na
    for f in (M.POS, M.VEL, M.MASS):
        np.testing.assert_array_equal(a.download(f, tot), b.download(f, tot))
    np.testing.assert_array_equal(ga.corresponding(tot), gb.corresponding(tot))
    c = O.Atoms.from_arrays(pos)
    og = O.Ghost()
    og.create_xyz(c, O.make_box([0, 0, 0], L, ghost))
    assert c.nlocal + c.nghost == tot
    np.testing.assert_array_equal(a.download(M.POS, tot), c.get(O.POS, tot))
    np.testing.assert_array_equal(ga.corresponding(tot), og.corr(tot))


def test_espp_verlet_pairs_sets_energy_forces(M, ctx, espp_positions):
    gbox, cbox, g, c = _espp(M, ctx, espp_positions)
    gl = M.GhostLayer(ctx)
    gl.create_ghost_atoms(g, gbox)
    og = O.Ghost()
    og.create_xyz(c, cbox)
    vl = M.HalfVerletList(ctx)
    vl.build(g, 0, g.nlocal, 2.8, 1.0, [-2.8] * 3, [L_ESPP + 2.8] * 3)
    assert vl.num_pairs() == 1310403
    ov = O.Verlet()
    ov.build(c, c.nlocal + c.nghost, 0, c.nlocal, 2.8, 1.0, [-2.8] * 3, [L_ESPP + 2.8] * 3)
    # half-list neighbour sets, every row
    gc, _ = vl.half_counts(g.nlocal)
    np.testing.assert_array_equal(gc, ov.counts()[: g.nlocal])
    grows = vl.half_rows(g.nlocal)
    onb = ov.neighbors()
    ocnt = ov.counts()
    for i in range(g.nlocal):
        assert grows[i] == set(onb[i, : ocnt[i]].tolist()), i
    # full rows are the union of both directions
    full, fcnt = vl.full_rows(g.nlocal)
    assert int(fcnt.sum()) > 2 * 1310403 - 22104 * 40
    # unshifted LJ energy of the ESPResSo++ comparison
    lj = M.LennardJones(ctx, 2.5, 1.0, 1.0)
    g.set_force(0.0)
    lj.apply(g, vl)
    gl.contribute_back_ghost_to_real(g)
    E, W = lj.energy_virial()
    assert abs(E - (-94795.927)) < 5e-4
    olj = O.LJ(0.0, 2.5, 1.0, 1.0)
    O.lib().or_set_threads(1)
    cE, cW = olj.apply(c, ov)
    og.fold(c)
    assert E == pytest.approx(cE, rel=1e-12)
    assert W == pytest.approx(cW, rel=1e-12)
    Fg = g.download(M.FORCE, g.nlocal)
    Fc = c.get(O.FORCE, c.nlocal)
    assert np.abs(Fg - Fc).max() <= F_TOL * np.abs(Fc).max()


@pytest.mark.parametrize("columns,cells,block", [(1, 1, 64), (1, 5, 256), (2, 2, 128), (2, 3, 256), (2, 4, 256),
                                                 (3, 1, 192)])
def test_tilings_give_identical_lists_and_forces(M, ctx, espp_positions, columns, cells, block):
    """the tile shape and workgroup size change neither the neighbour rows (set and order)
    nor a single bit of the forces: rows are summed in stencil order whatever the tiling"""
    gbox, cbox, g, c = _espp(M, ctx, espp_positions)
    gl = M.GhostLayer(ctx)
    gl.create_ghost_atoms(g, gbox)
    ref = M.HalfVerletList(ctx)
    ref.build(g, 0, g.nlocal, 2.8, 1.0, [-2.8] * 3, [L_ESPP + 2.8] * 3)
    vl = M.HalfVerletList(ctx)
    vl.set_tiling(cells, block, columns)
    vl.build(g, 0, g.nlocal, 2.8, 1.0, [-2.8] * 3, [L_ESPP + 2.8] * 3)
    t = vl.tiling()
    assert (t["tile_columns"], t["tile_cells"], t["block"]) == (columns, cells, block)
    assert vl.num_pairs() == 1310403
    a, ac = ref.full_rows(g.nlocal)
    b, bc = vl.full_rows(g.nlocal)
    np.testing.assert_array_equal(ac, bc)
    np.testing.assert_array_equal(a, b)
    lj = M.LennardJones(ctx, 2.5, 1.0, 1.0)
    g.set_force(0.0)
    lj.apply(g, ref)
    f_ref = g.download(M.FORCE, g.nlocal)
    g.set_force(0.0)
    lj.apply(g, vl)
    np.testing.assert_array_equal(g.download(M.FORCE, g.nlocal), f_ref)


def test_global_staging_equals_lds_staging(M, ctx, espp_positions):
    """tiles whose neighbourhood exceeds the LDS limit even at one cell column by one cell
    (dense inputs) stage it in global memory: forced here with a stage limit of 64 records,
    the rows (entries and order) and the forces equal those of LDS staging at the same tile
    shape bit for bit"""
    gbox, cbox, g, c = _espp(M, ctx, espp_positions)
    gl = M.GhostLayer(ctx)
    gl.create_ghost_atoms(g, gbox)
    lists = []
    for limit in (0, 64):
        vl = M.HalfVerletList(ctx)
        vl.set_tiling(1, 256, 1)
        vl.set_stage_limit(limit)
        vl.build(g, 0, g.nlocal, 2.8, 1.0, [-2.8] * 3, [L_ESPP + 2.8] * 3)
        t = vl.tiling()
        assert (t["tile_columns"], t["tile_cells"]) == (1, 1)
        lists.append(vl)
    assert lists[1].tiling()["max_staged"] > 64  # only global staging leaves a tile this large
    assert lists[1].num_pairs() == 1310403
    a, ac = lists[0].full_rows(g.nlocal)
    b, bc = lists[1].full_rows(g.nlocal)
    np.testing.assert_array_equal(ac, bc)
    np.testing.assert_array_equal(a, b)
    forces = []
    for vl in lists:
        for lj in (M.LennardJones(ctx, 2.5, 1.0, 1.0), M.LennardJones(ctx, 2.5, 1.0, 1.0, capping_distance=0.9)):
            g.set_force(0.0)
            lj.apply(g, vl)
            forces.append((g.download(M.FORCE, g.nlocal), lj.energy_virial()))
    for k in range(2):
        np.testing.assert_array_equal(forces[k][0], forces[2 + k][0])
        assert forces[k][1] == forces[2 + k][1]


def test_wide_offset_build_equals_narrow(M, ctx, espp_positions):
    """the 64-bit entry-offset form of the list kernel (which only lists of 2^31 entries or
    more, ~20M atoms, take by themselves) builds the same rows and gives the same forces"""
    gbox, cbox, g, c = _espp(M, ctx, espp_positions)
    gl = M.GhostLayer(ctx)
    gl.create_ghost_atoms(g, gbox)
    lists = []
    for wide in (False, True):
        vl = M.HalfVerletList(ctx)
        vl.set_wide_offsets(wide)
        vl.build(g, 0, g.nlocal, 2.8, 1.0, [-2.8] * 3, [L_ESPP + 2.8] * 3)
        lists.append(vl)
    a, ac = lists[0].full_rows(g.nlocal)
    b, bc = lists[1].full_rows(g.nlocal)
    np.testing.assert_array_equal(ac, bc)
    np.testing.assert_array_equal(a, b)
    lj = M.LennardJones(ctx, 2.5, 1.0, 1.0)
    forces = []
    for vl in lists:
        g.set_force(0.0)
        lj.apply(g, vl)
        forces.append(g.download(M.FORCE, g.nlocal))
    np.testing.assert_array_equal(forces[0], forces[1])


def test_espp_integrators_bit_exact(M, ctx, espp_positions):
    rng = np.random.default_rng(5)
    n = espp_positions.shape[0]
    vel = rng.standard_normal((n, 3))
    force = rng.standard_normal((n, 3)) * 10
    mass = 1.0 + rng.random(n)
    g, c = _both(M, ctx, espp_positions, vel=vel, force=force, mass=mass)
    d_g = M.VelocityVerlet.pre_force_integrate(g, 0.005)
    d_c = O.lib().or_vv_pre(c.h, 0.005)
    assert d_g == d_c
    np.testing.assert_array_equal(g.download(M.POS, n), c.get(O.POS, n))
    np.testing.assert_array_equal(g.download(M.VEL, n), c.get(O.VEL, n))
    M.VelocityVerlet.post_force_integrate(g, 0.005)
    O.lib().or_vv_post(c.h, 0.005)
    np.testing.assert_array_equal(g.download(M.VEL, n), c.get(O.VEL, n))
    # Langevin: counter-based noise gives identical bits on both sides
    lang = M.VelocityVerletLangevinThermostat(1.0, 1.2, seed=99)
    lang.call = 17
    d_g = lang.pre_force_integrate(g, 0.005)
    d_c = O.lib().or_langevin_pre(c.h, 0.005, 1.0, 1.2, 99, 17, None)
    assert d_g == d_c
    np.testing.assert_array_equal(g.download(M.VEL, n), c.get(O.VEL, n))
    np.testing.assert_array_equal(g.download(M.POS, n), c.get(O.POS, n))
    # predicate-restricted thermostat (symmetric slab around the box centre)
    p = M.pred1(M.PRED_SYMMETRIC_SLAB, 0, a=L_ESPP / 2, b=0.0, c=5.0)
    lang.call = 18
    lang.pre_force_integrate(g, 0.005, pred=p)
    op = O.pred1(kind=3, axis=0, a=L_ESPP / 2, b=0.0, c=5.0)
    O.lib().or_langevin_pre(c.h, 0.005, 1.0, 1.2, 99, 18, O.C.byref(op))
    np.testing.assert_array_equal(g.download(M.VEL, n), c.get(O.VEL, n))
    assert g.kinetic_energy() == pytest.approx(O.lib().or_kinetic_energy(c.h), rel=1e-13)


def test_cell_list_and_permute_bit_exact(M, ctx, espp_positions):
    g, c = _both(M, ctx, espp_positions, vel=np.arange(3 * 32768, dtype=float).reshape(-1, 3))
    delta = [2.8] * 3
    gcl = M.LinkedCellList(ctx, g, 0, g.nlocal, delta, [0, 0, 0], [L_ESPP] * 3)
    occ = O.Cells(c, 0, c.nlocal, delta, [0, 0, 0], [L_ESPP] * 3)
    for a, b in zip(gcl.download(), occ.get()):
        np.testing.assert_array_equal(a, b)
    n = g.nlocal
    idx = np.arange(n, dtype=float)
    g.upload(M.MASS, 1.0 + 1e-6 * idx)    # per-record values: moved with the records
    g.upload(M.CHARGE, -idx)
    g.fill(M.RELMASS, 0.5)                # one value for all: left where it is
    g.permute(gcl)
    occ.permute(c)
    np.testing.assert_array_equal(g.download(M.POS), c.get(O.POS))
    np.testing.assert_array_equal(g.download(M.VEL), c.get(O.VEL))
    src = g.download(M.VEL)[:, 0] / 3.0   # record k came from record src[k] (vel = arange)
    np.testing.assert_array_equal(g.download(M.MASS), 1.0 + 1e-6 * src)
    np.testing.assert_array_equal(g.download(M.CHARGE), -src)
    np.testing.assert_array_equal(g.download(M.RELMASS), np.full(n, 0.5))
    # a field that stops being uniform moves again on the next permutation (the same order
    # again: the records are already in cell order, so the permutation is the identity)
    g.upload(M.RELMASS, 2.0 + idx)
    gcl.build(g, 0, n, delta, [0, 0, 0], [L_ESPP] * 3)
    g.permute(gcl)
    np.testing.assert_array_equal(g.download(M.RELMASS), 2.0 + idx)
    np.testing.assert_array_equal(g.download(M.CHARGE), -src)


def test_ghost_update_and_fold(M, ctx, espp_positions):
    gbox, cbox, g, c = _espp(M, ctx, espp_positions)
    gl = M.GhostLayer(ctx)
    gl.create_ghost_atoms(g, gbox)
    og = O.Ghost()
    og.create_xyz(c, cbox)
    n = g.nlocal + g.nghost
    rng = np.random.default_rng(3)
    moved = c.get(O.POS, n)
    moved[: c.nlocal] += rng.uniform(-0.1, 0.1, (c.nlocal, 3))
    g.upload(M.POS, moved)
    c.set(O.POS, moved)
    gl.update_ghost_atoms(g, gbox)
    og.update(c, cbox)
    np.testing.assert_array_equal(g.download(M.POS, n), c.get(O.POS, n))
    f = rng.standard_normal((n, 3))
    g.upload(M.FORCE, f)
    c.set(O.FORCE, f)
    gl.contribute_back_ghost_to_real(g)
    og.fold(c)
    np.testing.assert_allclose(g.download(M.FORCE, n), c.get(O.FORCE, n), rtol=1e-13, atol=1e-13)


def _fcc_system(M, ctx, cells, rho=0.8442, T=0.72, seed=42):
    pos, L = fcc(cells, rho)
    vel = maxwell_boltzmann(pos.shape[0], T, seed)
    return pos, vel, L


@pytest.mark.parametrize("pred_kind", ["symmetric", "asymmetric"])
def test_lj_apply_if_predicates(M, ctx, pred_kind):
    pos, vel, L = _fcc_system(M, ctx, 6)
    pos = pos + np.random.default_rng(1).uniform(-0.15, 0.15, pos.shape)
    pos = np.mod(pos, L)
    g, c = _both(M, ctx, pos)
    gbox = M.make_box([0, 0, 0], [L] * 3, 2.8)
    cbox = O.make_box([0, 0, 0], [L] * 3, 2.8)
    gl = M.GhostLayer(ctx)
    gl.create_ghost_atoms(g, gbox)
    og = O.Ghost()
    og.create_xyz(c, cbox)
    vl = M.HalfVerletList(ctx)
    vl.build(g, 0, g.nlocal, 2.8, 1.0, [-2.8] * 3, [L + 2.8] * 3)
    ov = O.Verlet()
    ov.build(c, c.nlocal + c.nghost, 0, c.nlocal, 2.8, 1.0, [-2.8] * 3, [L + 2.8] * 3)
    region = dict(kind=M.PRED_OPEN_RANGE, axis=0, a=0.3 * L, b=0.6 * L)
    other = dict(kind=M.PRED_LESS, axis=1, a=0.5 * L)
    if pred_kind == "symmetric":
        gp = M.pred2(M.PAIR_OR, M.pred1(**region))
        cp = O.pred2(2, O.pred1(**region))
    else:
        gp = M.pred2(M.PAIR_AND, M.pred1(**region), M.pred1(**other))
        cp = O.pred2(1, O.pred1(**region), O.pred1(**other))
    lj = M.LennardJones(ctx, 2.5, 1.0, 1.0, capping_distance=0.9, shifted=True)
    olj = O.LJ(0.9, 2.5, 1.0, 1.0, 1, shifted=True)
    g.set_force(0.0)
    lj.apply(g, vl, gp)
    gl.contribute_back_ghost_to_real(g)
    cE, cW = olj.apply(c, ov, cp)
    og.fold(c)
    E, W = lj.energy_virial()
    assert E == pytest.approx(cE, rel=1e-12)
    assert W == pytest.approx(cW, rel=1e-12)
    Fg = g.download(M.FORCE, g.nlocal)
    Fc = c.get(O.FORCE, c.nlocal)
    assert np.abs(Fg - Fc).max() <= F_TOL * np.abs(Fc).max()


def test_lj_eval_and_table(M, ctx):
    lj = M.LennardJones(ctx, 2.5 * 3.01, 3.01, 2.0, capping_distance=0.1, shifted=True)
    olj = O.LJ(0.1, 2.5 * 3.01, 3.01, 2.0, 1, shifted=True)
    np.testing.assert_array_equal(lj.table(), olj.table())
    r = 0.2 + 0.1 * np.arange(100)
    ff, e = lj.eval(r * r)
    for k, x in enumerate(r):
        cff, ce = olj.eval(x * x)
        assert ff[k] == pytest.approx(cff, rel=1e-15, abs=0)
        assert e[k] == pytest.approx(ce, rel=1e-14, abs=1e-300)


def test_verlet_regrow_and_edge_cases(M, ctx):
    # tiny max_neigh forces the re-count path; empty and single-atom systems
    pos, _, L = _fcc_system(M, ctx, 3)
    g, c = _both(M, ctx, pos)
    gl = M.GhostLayer(ctx)
    gbox = M.make_box([0, 0, 0], [L] * 3, 2.8)
    gl.create_ghost_atoms(g, gbox)
    vl = M.HalfVerletList(ctx)
    vl.build(g, 0, g.nlocal, 2.8, 1.0, [-2.8] * 3, [L + 2.8] * 3, max_neigh=1)
    og = O.Ghost()
    og.create_xyz(c, O.make_box([0, 0, 0], [L] * 3, 2.8))
    ov = O.Verlet()
    ov.build(c, c.nlocal + c.nghost, 0, c.nlocal, 2.8, 1.0, [-2.8] * 3, [L + 2.8] * 3)
    assert vl.num_pairs() == ov.total()
    e = M.Atoms(ctx, 0)
    vl2 = M.HalfVerletList(ctx)
    vl2.build(e, 0, 0, 2.8, 1.0, [0, 0, 0], [10, 10, 10])
    assert vl2.num_pairs() == 0
    lj = M.LennardJones(ctx, 2.5, 1.0, 1.0)
    lj.apply(e, vl2)
    assert lj.energy_virial() == (0.0, 0.0)
    assert M.VelocityVerlet.pre_force_integrate(e, 0.005) == 0.0
    s = M.Atoms.from_arrays(ctx, np.array([[1.0, 1.0, 1.0]]))
    vl2.build(s, 0, 1, 2.8, 1.0, [0, 0, 0], [10, 10, 10])
    assert vl2.num_pairs() == 0


def test_invalid_arguments_fail_loudly(M, ctx):
    a = M.Atoms(ctx, 4)
    with pytest.raises(M.MrmdError):
        a.download(M.POS, 10)
    with pytest.raises(M.MrmdError):
        M.make_box  # noqa: B018 - attribute exists
        M.GhostLayer(ctx).create_ghost_atoms(a, M.make_box([0, 0, 0], [1, 1, 1], 2.0))
    bad = M.pred1(kind=99)
    with pytest.raises(M.MrmdError):
        M.VelocityVerletLangevinThermostat(1.0, 1.0).pre_force_integrate(a, 0.1, pred=bad)


def test_berendsen_matches_oracle(M, ctx):
    """BerendsenThermostat::apply (BerendsenThermostat.cpp:27-50): bit-exact velocity scaling
    against the oracle; the reference's single-atom cases (T = 41.5 -> 41.5 at gamma 0,
    -> 3.8 at gamma 1); T <= 0 leaves velocities alone; a target <= 0 is an error"""
    g = np.random.default_rng(8)
    pos, vel = g.uniform(0, 10, (5000, 3)), g.normal(size=(5000, 3))
    ga, ca = _both(M, ctx, pos, vel=vel)
    M.BerendsenThermostat.apply(ga, 1.3, 0.9, 0.2)
    assert O.lib().or_berendsen_apply(ca.h, 1.3, 0.9, 0.2) == 0
    np.testing.assert_array_equal(ga.download(M.VEL), ca.get(O.VEL))
    M.BerendsenThermostat.apply(ga, 0.0, 0.9, 0.2)
    np.testing.assert_array_equal(ga.download(M.VEL), ca.get(O.VEL))
    with pytest.raises(M.MrmdError):
        M.BerendsenThermostat.apply(ga, 1.0, 0.0, 0.2)
    for gamma, expected in ((0.0, 41.5), (1.0, 3.8)):
        a = M.Atoms.from_arrays(ctx, np.array([[2.0, 3.0, 4.0]]), vel=np.array([[7.0, 5.0, 3.0]]),
                                mass=np.array([1.5]))
        t = a.kinetic_energy() * (2.0 / 3.0)
        assert t == pytest.approx(41.5, rel=1e-15)
        M.BerendsenThermostat.apply(a, t, 3.8, gamma)
        assert a.kinetic_energy() * (2.0 / 3.0) == pytest.approx(expected, rel=1e-12)
