"""The AdResS LJ / ideal-gas kernels against the mathematics rather than the oracle: the
gradient identity (forces = -grad of the returned weighted energy, with the mean
compensation energy zero) and the literal compensation-energy fixture, on the tiled kernel
(k_igTiled: monatomic molecules, one type) and the half-list kernel (k_ljIdealGas: several
atoms per molecule or several types).  The CPU oracle is pinned by the same checks in
test_oracle_adress_pins.py (action/LJ_IdealGas.cpp:21-50, 52-215).
"""
import math

import numpy as np
import pytest

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


def _jittered_grid(shape, spacing, jitter, seed):
    g = np.random.default_rng(seed)
    idx = np.stack(np.meshgrid(*[np.arange(s) for s in shape], indexing="ij"), -1).reshape(-1, 3)
    return (idx + 0.5) * spacing + g.uniform(-jitter, jitter, idx.shape)


def _forces(M, ga, gm, vl, lj, slab, gl=None, box=None):
    if gl is not None:
        gl.update_ghost_atoms(ga, box)
    gm.update(ga, slab)
    ga.set_force(0.0)
    gm.set_force(0.0)
    lj.set_histogram(2, np.zeros((200, lj.num_types)))
    e = lj.run(gm, vl, ga)
    gm.contribute_force(ga)
    if gl is not None:
        gl.contribute_back_ghost_to_real(ga)
    return e, ga.download(M.FORCE, ga.nlocal)


def _fd(M, ga, gm, vl, lj, slab, rows, gl=None, box=None, h=1e-6):
    _, F = _forces(M, ga, gm, vl, lj, slab, gl, box)
    n = ga.nlocal
    pos = ga.download(M.POS, n)
    fd = np.zeros((len(rows), 3))
    for k, i in enumerate(rows):
        for d in range(3):
            for sgn in (1.0, -1.0):
                p = pos[i].copy()
                p[d] += sgn * h
                ga.upload(M.POS, p.reshape(1, 3), offset=int(i))
                e, _ = _forces(M, ga, gm, vl, lj, slab, gl, box)
                fd[k, d] -= sgn * e / (2.0 * h)
            ga.upload(M.POS, pos[i].reshape(1, 3), offset=int(i))
    err = np.abs(F[rows] - fd) / (1.0 + np.abs(F[rows]))
    assert err.max() <= 1e-6, err.max()
    return F[rows]


@pytest.mark.parametrize("kernel,num_types", [("lj_idealgas_tiled", 1), ("lj_idealgas", 2)])
def test_gradient_identity_periodic_slab_across_boundary(M, ctx, kernel, num_types):
    """periodic 24 x 7.2 x 7.2 box, 1,440 monatomic molecules, multi-resolution ghosts, slab
    (nu = 2) centred at x = 2 so that hybrid ghost images face coarse-grained reals; one type
    runs the tiled kernel, two types (a 2 x 2 table) the half-list kernel"""
    L = [24.0, 7.2, 7.2]
    rn = 2.8
    pos = np.mod(_jittered_grid((20, 6, 6), 1.2, 0.15, 8), L)
    types = (np.random.default_rng(2).random(pos.shape[0]) < 0.3).astype(np.int64) if num_types == 2 else None
    ga = M.Atoms.from_arrays(ctx, pos, type=types)
    gm = M.Molecules.for_each_atom(ctx, ga)
    box = M.make_box([0, 0, 0], L, rn)
    gl = M.GhostLayer(ctx)
    gm.update(ga, None)
    gl.multires_create_ghost_atoms(gm, ga, box)
    gm.update(ga, None)
    vl = M.HalfVerletList(ctx)
    vl.build(gm, 0, gm.nlocal, rn, 1.0, [-rn] * 3, [L[0] + rn, L[1] + rn, L[2] + rn], molecules=True)
    slab = M.Slab([2.0, 0, 0], 6.0, 5.0, 2)
    if num_types == 2:
        lj = M.LJ_IdealGas(ctx, [0.0] * 4, [2.5] * 4, [1.0, 0.95, 0.95, 0.9], [1.0, 0.8, 0.8, 0.6], True,
                           num_types=2)
    else:
        lj = M.LJ_IdealGas(ctx, [0.0], [2.5], [1.0], [1.0], True)
    x = pos[:, 0]
    dx = np.minimum(np.abs(x - 2.0), L[0] - np.abs(x - 2.0))
    hyb = np.where((dx > 3.001) & (dx < 7.999))[0]
    rows = list(hyb[:: max(1, len(hyb) // 30)][:30]) + list(np.where(dx < 2.9)[0][:3])
    assert any(x[r] > 16.0 for r in rows) and any(x[r] < 8.0 for r in rows)
    ctx.set_profiling(True)
    ctx.reset_stats()
    _fd(M, ga, gm, vl, lj, slab, rows, gl, box)
    launches = ctx.kernel_stats(kernel)[0]
    ctx.set_profiling(False)
    assert launches == 1 + 6 * len(rows)
    lam = gm.download(M.M_MODLAMBDA, gm.nlocal + gm.nghost)
    assert (lam[gm.nlocal:] > 0).any()


def test_gradient_identity_diatomic_open(M, ctx):
    """open box, 400 diatomic molecules (relMass 1/2, bond 0.4 sigma): the half-list
    kernel with molecule forces contributed to two atoms each"""
    com = _jittered_grid((16, 5, 5), 1.5, 0.1, 3)
    g = np.random.default_rng(4)
    u = g.normal(size=com.shape)
    u /= np.linalg.norm(u, axis=1, keepdims=True)
    pos = np.empty((2 * com.shape[0], 3))
    pos[0::2] = com + 0.2 * u
    pos[1::2] = com - 0.2 * u
    nm = com.shape[0]
    ga = M.Atoms.from_arrays(ctx, pos, relmass=np.full(pos.shape[0], 0.5))
    gm = M.Molecules(ctx, nm)
    gm.upload(M.M_OFFSET, np.arange(nm) * 2)
    gm.upload(M.M_NUMATOMS, np.full(nm, 2))
    gm.set_counts(nm, 0)
    slab = M.Slab([12.0, 0, 0], 6.0, 5.0, 2)
    gm.update(ga, slab)
    vl = M.HalfVerletList(ctx)
    vl.build(gm, 0, nm, 2.5 + 0.4 + 0.3, 1.0, [-1.0] * 3, [25.0, 8.5, 8.5], molecules=True)
    lj = M.LJ_IdealGas(ctx, [0.0], [2.5], [1.0], [1.0], True)
    cx = gm.download(M.M_POS, nm)[:, 0]
    ax = np.abs(cx - 12.0)
    hyb = np.where((ax > 3.001) & (ax < 7.999))[0]
    mrows = list(hyb[:: max(1, len(hyb) // 20)][:20])
    rows = [2 * m for m in mrows] + [2 * m + 1 for m in mrows[:10]]
    ctx.set_profiling(True)
    ctx.reset_stats()
    _fd(M, ga, gm, vl, lj, slab, rows)
    assert ctx.kernel_stats("lj_idealgas")[0] == 1 + 6 * len(rows)
    ctx.set_profiling(False)


def test_compensation_energy_triangle_literal(M, ctx):
    """three hybrid monatomic molecules on an equilateral triangle of side 2^(1/6), LJ
    epsilon = sigma = 1 unshifted: every pair energy is -1, so each molecule's bin of lambda
    collects -1 per sampling run; with sampling every run and an update every second run
    the mean is -1/11 after run 0 and -21/121 after run 2, and the molecule forces are
    grad(lambda_mod) (drift) plus mean x grad(lambda_mod)"""
    s = 2.0 ** (1.0 / 6.0)
    pos = np.array([[0.0, 0.0, 0.0], [s, 0.0, 0.0], [0.5 * s, 0.5 * math.sqrt(3.0) * s, 0.0]])
    lam = [0.3013, 0.5527, 0.8061]
    mlam = [0.4, 0.6, 0.7]
    grad = np.array([[-0.5, 0.25, 0.0], [0.75, 0.0, -0.125], [0.0, -1.5, 0.5]])
    gm = M.Molecules(ctx, 3)
    gm.upload(M.M_POS, pos)
    gm.upload(M.M_OFFSET, [0, 1, 2])
    gm.upload(M.M_NUMATOMS, [1, 1, 1])
    gm.upload(M.M_LAMBDA, lam)
    gm.upload(M.M_MODLAMBDA, mlam)
    gm.upload(M.M_GRAD, grad)
    gm.set_counts(3, 0)
    ga = M.Atoms.from_arrays(ctx, pos)
    vl = M.HalfVerletList(ctx)
    vl.build(gm, 0, 3, 2.0, 1.0, [-2.0] * 3, [3.0] * 3, molecules=True)
    assert vl.num_pairs() == 3
    lj = M.LJ_IdealGas(ctx, [0.0], [2.5], [1.0], [1.0], False)
    lj.set_intervals(1, 2)
    bins = [60, 110, 161]

    def run():
        ga.set_force(0.0)
        gm.set_force(0.0)
        e = lj.run(gm, vl, ga)
        return e, gm.download(M.M_FORCE, 3)

    e, Fm = run()
    w = [(mlam[0] + mlam[1]) / 2, (mlam[0] + mlam[2]) / 2, (mlam[1] + mlam[2]) / 2]
    assert e == pytest.approx(-sum(w), abs=1e-14)
    np.testing.assert_allclose(Fm, grad, atol=1e-14)
    assert np.abs(ga.download(M.FORCE, 3)).max() < 1e-12
    m1 = np.zeros((200, 1))
    m1[bins, 0] = -1.0 / 11.0
    np.testing.assert_allclose(lj.histogram(2), m1, atol=1e-15)
    e, Fm = run()
    np.testing.assert_allclose(Fm, grad * (1.0 - 1.0 / 11.0), atol=1e-14)
    c = np.zeros((200, 1))
    c[bins, 0] = -1.0
    np.testing.assert_allclose(lj.histogram(0), c, atol=1e-14)
    cnt = np.zeros((200, 1))
    cnt[bins, 0] = 1.0
    np.testing.assert_array_equal(lj.histogram(1), cnt)
    run()
    m2 = np.zeros((200, 1))
    m2[bins, 0] = -21.0 / 121.0
    np.testing.assert_allclose(lj.histogram(2), m2, atol=1e-15)
    assert not lj.histogram(0).any() and not lj.histogram(1).any()
