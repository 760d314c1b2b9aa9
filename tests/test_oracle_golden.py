"""Pins the CPU oracle against the reference's own golden vectors and known answers.

Each test names the reference test it restates.  None of these needs a GPU.
"""
import math

import numpy as np
import pytest

import pyoracle as O


@pytest.fixture(autouse=True, scope="module")
def _serial():
    O.lib().or_set_threads(1)
    yield
    O.lib().or_set_threads(4)


# ---------------------------------------------------------------- ESPResSo++ comparison
# tests/LennardJones/LennardJones.cpp:30-38, 82-127
ESPP_L = 33.8585


def _espp_system(pos):
    box = O.make_box([0, 0, 0], [ESPP_L] * 3, 2.8)
    a = O.Atoms.from_arrays(pos)
    return box, a


def test_espp_ghosts_pairs_energy(espp_positions):
    O.lib().or_set_threads(4)
    box, a = _espp_system(espp_positions)
    g = O.Ghost()
    g.create_xyz(a, box)
    assert a.nlocal == 32768
    assert a.nghost == 22104
    vl = O.Verlet()
    vl.build(a, a.nlocal + a.nghost, 0, a.nlocal, 2.8, 1.0, [-2.8] * 3, [ESPP_L + 2.8] * 3)
    assert vl.total() == 1310403
    lj = O.LJ(0.0, 2.5, 1.0, 1.0)  # LennardJones(rc, sigma, epsilon): unshifted, uncapped
    E, W = lj.apply(a, vl)
    # EXPECT_FLOAT_EQ semantics (4 ulp in float) and a tighter check on the stored digits
    assert abs(E - (-94795.927)) < 5e-4
    O.lib().or_set_threads(1)


def _brute_count(pos, cutoff, box, periodic):
    # countWithinCutoff of tests/LennardJones/LennardJones.cpp:40-70 (vectorised)
    n = pos.shape[0]
    count = 0
    for i in range(n):
        d = np.abs(pos[i + 1:] - pos[i])
        if periodic:
            d = np.where(d > box * 0.5, d - box, d)
        count += int(np.sum(np.sum(d * d, axis=1) < cutoff * cutoff))
    return count


def test_espp_brute_force_counts(espp_positions):
    # periodic brute force = 1,310,403 ; non-periodic over local+ghost = 1,426,948
    sub = espp_positions[::1]
    assert _brute_count(sub, 2.8, ESPP_L, True) == 1310403
    box, a = _espp_system(espp_positions)
    g = O.Ghost()
    g.create_xyz(a, box)
    allpos = a.get(O.POS)
    # non-periodic count over pairs (i local, j > i in local+ghost)
    cnt = 0
    n = a.nlocal
    for i in range(n):
        d = allpos[i + 1:] - allpos[i]
        cnt += int(np.sum(np.sum(d * d, axis=1) < 2.8 * 2.8))
    assert cnt == 1426948


def test_half_list_equals_brute_force_sets(espp_positions):
    """every half-list row equals the brute-force discriminator set (first 2,000 rows)"""
    box, a = _espp_system(espp_positions)
    g = O.Ghost()
    g.create_xyz(a, box)
    vl = O.Verlet()
    vl.build(a, a.nlocal + a.nghost, 0, a.nlocal, 2.8, 1.0, [-2.8] * 3, [ESPP_L + 2.8] * 3)
    P = a.get(O.POS)
    rows = vl.rows(0, 2000)
    for i in range(2000):
        d = P[i] - P
        d2 = d[:, 0] * d[:, 0] + d[:, 1] * d[:, 1] + d[:, 2] * d[:, 2]
        xp, xn = P[i], P
        lex = (xn[:, 0] > xp[0]) | ((xn[:, 0] == xp[0]) & ((xn[:, 1] > xp[1]) |
                                                           ((xn[:, 1] == xp[1]) & (xn[:, 2] > xp[2]))))
        ok = (d2 <= 2.8 * 2.8) & lex
        ok[i] = False
        assert rows[i] == set(np.nonzero(ok)[0].tolist()), i


# ---------------------------------------------------------------- LJ closed form
# mrmd/action/LennardJones.test.cpp:46-81
def test_lj_explicit_comparison():
    eps, sigma = 2.0, 3.01
    rc = 2.5 * sigma
    cap = 0.1
    cutoff_pot = 4 * eps * ((sigma / rc) ** 12 - (sigma / rc) ** 6)
    lj = O.LJ(cap, rc, sigma, eps, 1, shifted=True)
    for idx in range(100):
        x = cap + 0.1 + idx * 0.1
        ff, e = lj.eval(x * x)
        pot = 4 * eps * ((sigma / x) ** 12 - (sigma / x) ** 6) - cutoff_pot
        force = 4 * eps * (-12 * (sigma / x) ** 12 + 6 * (sigma / x) ** 6) * x / (x * x)
        assert e == pytest.approx(pot, rel=1e-12, abs=1e-12)
        assert -x * ff == pytest.approx(force, rel=1e-12, abs=1e-12)


def test_lj_capping_continuity():
    """below the capping distance the force is constant cc/r * r and the energy linear; both
    continuous at r_cap (LennardJones.hpp:70-77, LennardJones.cpp:61-79)"""
    lj = O.LJ(0.9, 2.5, 1.0, 1.0, 1, shifted=True)
    rcap = 0.9
    ff_in, e_in = lj.eval((rcap - 1e-9) ** 2)
    ff_out, e_out = lj.eval((rcap + 1e-9) ** 2)
    assert ff_in * (rcap - 1e-9) == pytest.approx(ff_out * (rcap + 1e-9), rel=1e-6)
    assert e_in == pytest.approx(e_out, rel=1e-6)
    # linear energy inside: E(r) = E_cap - (r - r_cap) cc - shift
    f1, e1 = lj.eval(0.5 ** 2)
    f2, e2 = lj.eval(0.6 ** 2)
    slope = (e2 - e1) / 0.1
    assert slope == pytest.approx(-(f1 * 0.5), rel=1e-9)


def test_lj_force_is_minus_gradient():
    lj = O.LJ(0.0, 2.5, 1.0, 1.0, 1, shifted=True)
    for r in np.linspace(0.95, 2.45, 31):
        h = 1e-6
        _, ep = lj.eval((r + h) ** 2)
        _, em = lj.eval((r - h) ** 2)
        ff, _ = lj.eval(r * r)
        # F(r) along d = ff * r ; -dE/dr = ff * r
        assert -(ep - em) / (2 * h) == pytest.approx(ff * r, rel=1e-6, abs=1e-8)


def test_lj_newton_third_law_and_virial(espp_positions):
    box, a = _espp_system(espp_positions[:4000] * 0.5)
    box = O.make_box([0, 0, 0], [ESPP_L * 0.5] * 3, 2.8)
    a = O.Atoms.from_arrays(np.mod(espp_positions[:4000], ESPP_L * 0.5))
    g = O.Ghost()
    g.create_xyz(a, box)
    vl = O.Verlet()
    vl.build(a, a.nlocal + a.nghost, 0, a.nlocal, 2.8, 1.0, [-2.8] * 3, [ESPP_L * 0.5 + 2.8] * 3)
    lj = O.LJ(0.8, 2.5, 1.0, 1.0, 1, shifted=True)
    lj.apply(a, vl)
    g.fold(a)
    F = a.get(O.FORCE)[: a.nlocal]
    assert np.abs(F.sum(axis=0)).max() < 1e-9 * np.abs(F).max() * a.nlocal ** 0.5


# ---------------------------------------------------------------- ghost exchange
# mrmd/communication/GhostExchange.test.cpp:64-153
def _grid27():
    box = O.make_box([0, 0, 0], [3, 3, 3], 0.7)
    pts = [(x + 0.5, y + 0.5, z + 0.5) for x in range(3) for y in range(3) for z in range(3)]
    return box, O.Atoms.from_arrays(np.array(pts))


@pytest.mark.parametrize("axis", [0, 1, 2])
def test_ghost_self_exchange_axis(axis):
    box, a = _grid27()
    g = O.Ghost()
    g.reset(a)
    g.create_axis(a, box, axis)
    assert a.nghost == 18
    corr = g.corr(a.nlocal + a.nghost)
    assert (corr[:27] == -1).all()
    assert (corr[27:] < 27).all() and (corr[27:] >= 0).all()


def test_ghost_self_exchange_xyz_and_pairs():
    box, a = _grid27()
    g = O.Ghost()
    g.create_xyz(a, box)
    assert a.nghost == 98
    P = a.get(O.POS)
    cnt = 0
    for i in range(27):
        d = P[i + 1:] - P[i]
        cnt += int(np.sum(np.sum(d * d, axis=1) < 1.1 * 1.1))
    assert cnt == 108
    # the ghost images sit exactly at real position +- L per axis
    corr = g.corr(a.nlocal + a.nghost)
    for gi in range(27, 27 + 98):
        shift = P[gi] - P[corr[gi]]
        assert np.all(np.isclose(np.abs(shift), 0) | np.isclose(np.abs(shift), 3.0))


# mrmd/communication/PeriodicMapping.test.cpp:46-87
@pytest.mark.parametrize("initial,mapped", [
    ((0.4, 0.5, 0.6), (0.4, 0.5, 0.6)),
    ((1.1, 0.5, 0.6), (0.1, 0.5, 0.6)), ((-0.1, 0.5, 0.6), (0.9, 0.5, 0.6)),
    ((0.4, 1.1, 0.6), (0.4, 0.1, 0.6)), ((0.4, -0.1, 0.6), (0.4, 0.9, 0.6)),
    ((0.4, 0.5, 1.1), (0.4, 0.5, 0.1)), ((0.4, 0.5, -0.1), (0.4, 0.5, 0.9)),
    ((1.1, 1.2, 1.3), (0.1, 0.2, 0.3)), ((-0.3, -0.2, -0.1), (0.7, 0.8, 0.9))])
def test_periodic_mapping(initial, mapped):
    box = O.make_box([0, 0, 0], [1, 1, 1], 0.0)
    a = O.Atoms.from_arrays(np.array([initial]))
    O.lib().or_map_into_domain(a.h, O.C.byref(box))
    np.testing.assert_allclose(a.get(O.POS)[0], mapped, rtol=1e-7)


# mrmd/communication/UpdateGhostAtoms.test.cpp:69-109
@pytest.mark.parametrize("delta,final", [
    ((0.2, 0, 0), (1, 0, 0)), ((-0.2, 0, 0), (-1, 0, 0)), ((0, 0.2, 0), (0, 1, 0)),
    ((0, -0.2, 0), (0, -1, 0)), ((0, 0, 0.2), (0, 0, 1)), ((0, 0, -0.2), (0, 0, -1)),
    ((0.2, 0.2, 0.2), (1, 1, 1)), ((-0.2, -0.2, -0.2), (-1, -1, -1))])
def test_update_ghost_atoms(delta, final):
    box = O.make_box([0, 0, 0], [1, 1, 1], 0.1)
    a = O.Atoms.from_arrays(np.array([[0.5, 0.5, 0.5], np.add([0.5] * 3, delta)]), nghost=1)
    g = O.Ghost()
    g.set_corr([-1, 0])
    g.update(a, box)
    np.testing.assert_allclose(a.get(O.POS)[1], np.add([0.5] * 3, final), rtol=1e-12)


# mrmd/communication/AccumulateForce.test.cpp:27-49
def test_accumulate_force():
    a = O.Atoms.from_arrays(np.zeros((101, 3)), force=np.ones((101, 3)), nghost=100)
    g = O.Ghost()
    g.set_corr([-1] + [0] * 100)
    g.fold(a)
    np.testing.assert_allclose(a.get(O.FORCE)[0], [101, 101, 101])
    assert np.all(a.get(O.FORCE)[1:] == 0)


# ---------------------------------------------------------------- integrators
# mrmd/action/VelocityVerlet.test.cpp:27-71, mrmd/test/SingleAtom.hpp:28-60
def _single_atom():
    a = O.Atoms.from_arrays(np.array([[2.0, 3.0, 4.0]]), vel=np.array([[7.0, 5.0, 3.0]]),
                            force=np.array([[9.0, 7.0, 8.0]]), mass=np.array([1.5]))
    a.set(O.CHARGE, [0.5])
    return a


def test_vv_pre():
    a = _single_atom()
    O.lib().or_vv_pre(a.h, 4.0)
    np.testing.assert_allclose(a.get(O.FORCE)[0], [9, 7, 8])
    np.testing.assert_allclose(a.get(O.VEL)[0], [19, 14.333333, 13.666667], rtol=1e-7)
    np.testing.assert_allclose(a.get(O.POS)[0], [78, 60.333332, 58.666668], rtol=1e-7)


def test_vv_post():
    a = _single_atom()
    O.lib().or_vv_post(a.h, 4.0)
    np.testing.assert_allclose(a.get(O.VEL)[0], [19, 14.333333, 13.666667], rtol=1e-7)
    np.testing.assert_allclose(a.get(O.POS)[0], [2, 3, 4])


# mrmd/action/VelocityVerletLangevinThermostat.test.cpp:44-121
def test_langevin_no_thermostat_equals_vv():
    a = _single_atom()
    never = O.pred1(kind=1)
    O.lib().or_langevin_pre(a.h, 4.0, 0.5, 0.5, 1234, 0, O.C.byref(never))
    np.testing.assert_allclose(a.get(O.VEL)[0], [19, 14.333333, 13.666667], rtol=1e-7)
    np.testing.assert_allclose(a.get(O.POS)[0], [78, 60.333332, 58.666668], rtol=1e-7)


def test_langevin_simple_differs():
    a = _single_atom()
    O.lib().or_langevin_pre(a.h, 4.0, 0.5, 0.5, 1234, 0, None)
    v = a.get(O.VEL)[0]
    assert not np.allclose(v, [19, 14.333333, 13.666667], atol=1e-6)


def test_langevin_local_thermostat():
    a = _single_atom()
    local = O.pred1(kind=7, axis=0, a=78.0)  # x > 78
    O.lib().or_langevin_pre(a.h, 4.0, 0.5, 0.5, 1234, 0, O.C.byref(local))
    np.testing.assert_allclose(a.get(O.POS)[0], [78, 60.333332, 58.666668], rtol=1e-7)
    O.lib().or_langevin_pre(a.h, 4.0, 0.5, 0.5, 1234, 1, O.C.byref(local))
    assert not np.allclose(a.get(O.VEL)[0], [31, 23.666667, 24.333333], atol=1e-6)


def test_langevin_temperature_relaxation():
    """tests/LangevinThermostat/LangevinThermostat.cpp:82-99: 1e5 free atoms, zeta=1e5,
    dt=1e-3, 21 steps -> T = 1.12 +- 0.01 (here with the counter-based noise)"""
    O.lib().or_set_threads(4)
    rng = np.random.default_rng(1234)
    n = 100000
    pos = rng.random((n, 3)) * 10
    vel = (rng.random((n, 3)) - 0.5) * 10
    a = O.Atoms.from_arrays(pos, vel=vel)
    for step in range(21):
        O.lib().or_langevin_pre(a.h, 1e-3, 1e5, 1.12, 1234, step, None)
        O.lib().or_vv_post(a.h, 1e-3)
    T = (2.0 / 3.0) * O.lib().or_kinetic_energy(a.h) / n
    assert abs(T - 1.12) < 0.01
    O.lib().or_set_threads(1)


# ---------------------------------------------------------------- RNG / deterministic math
def test_philox_known_answers():
    # Random123 kat_vectors for philox4x32_10
    cases = [
        ([0, 0, 0, 0], [0, 0], [0x6627e8d5, 0xe169c58d, 0xbc57ac4c, 0x9b00dbd8]),
        ([0xffffffff] * 4, [0xffffffff] * 2, [0x408f276d, 0x41c83b0e, 0xa20bc7c6, 0x6d5451fd]),
        ([0x243f6a88, 0x85a308d3, 0x13198a2e, 0x03707344], [0xa4093822, 0x299f31d0],
         [0xd16cfe09, 0x94fdcceb, 0x5001e420, 0x24126ea1]),
    ]
    for ctr, key, exp in cases:
        c = np.array(ctr, dtype=np.uint32)
        k = np.array(key, dtype=np.uint32)
        out = np.zeros(4, dtype=np.uint32)
        O.lib().or_philox4x32_10(O.ptr(c), O.ptr(k), O.ptr(out))
        assert [int(x) for x in out] == exp


def test_det_log_exp_accuracy():
    xs = np.concatenate([np.logspace(-300, 0, 2000), np.random.default_rng(0).random(2000)])
    for x in xs:
        if x <= 0:
            continue
        assert O.lib().or_det_log(x) == pytest.approx(math.log(x), rel=2e-15, abs=1e-15)
    for x in np.concatenate([-np.logspace(-12, 2.8, 2000), [0.0]]):
        assert O.lib().or_det_exp(x) == pytest.approx(math.exp(x), rel=4e-15)


def test_normals_moments():
    out = np.zeros(3)
    vals = []
    for i in range(40000):
        O.lib().or_normals3(1234, i, 7, O.ptr(out))
        vals.append(out.copy())
    v = np.array(vals).ravel()
    assert abs(v.mean()) < 0.01
    assert abs(v.var() - 1.0) < 0.01
    assert abs(np.mean(v ** 4) - 3.0) < 0.06


# ---------------------------------------------------------------- AdResS LJ / ideal gas
# mrmd/action/LJ_IdealGas.test.cpp:34-220
def _adress_fixture(modlambda, force0=0.0):
    mols = O.Molecules(2)
    mols.set(O.M_POS, [[-0.5, 0, 0], [0.5, 0, 0]])
    mols.set(O.M_OFFSET, [0, 2])
    mols.set(O.M_NUMATOMS, [2, 2])
    mols.set(O.M_MODLAMBDA, [modlambda, modlambda])
    mols.set_counts(2, 0)
    pos = np.array([[-0.5, -0.5, 0], [-0.5, 0.5, 0], [0.5, -0.5, 0], [0.5, 0.5, 0]])
    atoms = O.Atoms.from_arrays(pos, force=np.full((4, 3), force0), relmass=np.full(4, 0.5))
    vl = O.Verlet()
    vl.build(mols, 2, 0, 2, 2.0, 1.0, [-1, -1, -1], [1, 1, 1], 4, molecules=True)
    return mols, atoms, vl


@pytest.mark.parametrize("lam,scale", [(1.0, 1.0), (0.5, 0.5)])
def test_lj_ideal_gas_forces(lam, scale):
    mols, atoms, vl = _adress_fixture(lam)
    sigma = 0.9
    lj = O.LJIdealGas(0.0, 2.5 * sigma, sigma, 2.0, 1, shifted=True)
    lj.run(mols, vl, atoms)
    F = atoms.get(O.FORCE)
    fx, fy = 0.22156665 * scale, 1.3825009 * scale
    expect = [[-fx, +fy, 0], [-fx, -fy, 0], [+fx, +fy, 0], [+fx, -fy, 0]]
    np.testing.assert_allclose(F, expect, rtol=2e-7, atol=1e-12)


def test_lj_ideal_gas_cg_untouched():
    mols, atoms, vl = _adress_fixture(0.0, force0=2.0)
    lj = O.LJIdealGas(0.0, 2.25, 0.9, 2.0, 1, shifted=True)
    lj.run(mols, vl, atoms)
    assert np.all(atoms.get(O.FORCE) == 2.0)


# ---------------------------------------------------------------- thermodynamic force
# mrmd/action/ThermodynamicForce.test.cpp:96-245
def _tf():
    box = O.make_box([0, 0, 0], [10, 1, 1], 1.0)
    tf = O.ThermodynamicForce([1.0], box, 1.0, [1.0])
    tf.set_force(np.arange(tf.nbins, dtype=float).reshape(-1, 1))
    return tf


def _atoms_uniform():
    return O.Atoms.from_arrays(np.array([[i / 10.0, 0, 0] for i in range(100)]))


def _atoms_nonuniform():
    pts = [[i + 0.5, 0, 0] for i in range(10) for _ in range(i)]
    return O.Atoms.from_arrays(np.array(pts))


def test_tf_sample():
    tf = _tf()
    tf.sample(_atoms_uniform())
    assert np.all(tf.density()[:, 0] == 10.0)
    tf2 = _tf()
    tf2.sample(_atoms_nonuniform())
    np.testing.assert_array_equal(tf2.density()[:, 0], np.arange(10.0))


def test_tf_apply_and_apply_if():
    a = _atoms_uniform()
    _tf().apply(a)
    x = a.get(O.POS)[:, 0]
    np.testing.assert_allclose(a.get(O.FORCE)[:, 0], np.floor(x))
    a = _atoms_uniform()
    _tf().apply(a, O.pred1(kind=2, a=0.5, b=4.5))
    inside = (x > 0.5) & (x < 4.5)
    np.testing.assert_allclose(a.get(O.FORCE)[:, 0], np.where(inside, np.floor(x), 0.0))


def test_tf_apply_interpolated_if():
    a = _atoms_uniform()
    _tf().apply_interpolated(a, O.pred1(kind=2, a=0.5, b=4.5))
    x = a.get(O.POS)[:, 0]
    inside = (x > 0.5) & (x < 4.5)
    np.testing.assert_allclose(a.get(O.FORCE)[:, 0], np.where(inside, x - 0.5, 0.0), atol=1e-12)


def test_tf_update_and_update_if():
    tf = _tf()
    tf.sample(_atoms_nonuniform())
    assert tf.update(1.0, 0.0) == 0
    np.testing.assert_allclose(tf.force()[:, 0], np.arange(10.0) - 1.0)
    assert tf.samples == 0 and np.all(tf.density() == 0)
    tf = _tf()
    tf.sample(_atoms_nonuniform())
    tf.update(1.0, 0.0, O.pred1(kind=2, a=0.51, b=4.49))
    grid = np.arange(10) + 0.5
    inside = (grid > 0.51) & (grid < 4.49)
    np.testing.assert_allclose(tf.force()[:, 0], np.where(inside, np.arange(10.0) - 1, np.arange(10.0)))


def test_tf_update_without_samples_fails():
    assert _tf().update(1.0, 0.0) == -1


# ---------------------------------------------------------------- MultiHistogram
# mrmd/data/MultiHistogram.test.cpp:25-293
def test_hist_get_bin_and_position():
    L = O.lib()
    assert L.or_hist_get_bin(0.0, 10.0, 10, -0.5) == -1
    assert L.or_hist_get_bin(0.0, 10.0, 10, 0.5) == 0
    assert L.or_hist_get_bin(0.0, 10.0, 10, 5.5) == 5
    assert L.or_hist_get_bin(0.0, 10.0, 10, 10.5) == -1
    assert L.or_hist_bin_position(0.0, 10.0, 10, 0) == 0.5
    assert L.or_hist_bin_position(0.0, 10.0, 10, 5) == 5.5
    for b in range(10):
        assert L.or_hist_get_bin(0.0, 10.0, 10, L.or_hist_bin_position(0.0, 10.0, 10, b)) == b


def test_hist_make_symmetric():
    d = np.zeros((10, 2))
    d[:, 0] = np.arange(10)
    d[:, 1] = 10 - np.arange(10)
    O.lib().or_hist_make_symmetric(10, 2, O.ptr(d))
    assert np.allclose(d[:, 0], 4.5) and np.allclose(d[:, 1], 5.5)


def test_hist_gradient():
    d = np.zeros((10, 3))
    d[:, 0] = 1
    d[:, 1] = np.arange(10)
    d[:, 2] = 10 - np.arange(10)
    out = np.zeros_like(d)
    O.lib().or_hist_gradient(0.0, 10.0, 10, 3, O.ptr(d), 0, O.ptr(out))
    np.testing.assert_allclose(out, np.tile([0.0, 1.0, -1.0], (10, 1)))
    # periodic: interior unchanged, ends wrap
    O.lib().or_hist_gradient(0.0, 10.0, 10, 3, O.ptr(d), 1, O.ptr(out))
    assert out[0, 1] == pytest.approx((1 - 9) * 0.5)


def test_hist_smoothen():
    d = np.zeros((11, 2))
    d[5] = [10, 5]
    out = np.zeros_like(d)
    O.lib().or_hist_smoothen(0.0, 10.0, 11, 2, O.ptr(d), 1.0, 3.0, 0, O.ptr(out))
    for i in range(1, 6):
        np.testing.assert_allclose(out[5 - i], out[5 + i], rtol=1e-7)
    c = np.full((11, 1), 2.0)
    out = np.zeros_like(c)
    O.lib().or_hist_smoothen(0.0, 10.0, 11, 1, O.ptr(c), 1.0, 3.0, 0, O.ptr(out))
    np.testing.assert_allclose(out, 2.0, rtol=1e-12)
    # closed form of the Gaussian kernel weights for an impulse (non-periodic, centre bin)
    bs = 10.0 / 11
    w = np.exp(-(np.arange(-3, 4) * bs) ** 2)
    assert out is not None
    O.lib().or_hist_smoothen(0.0, 10.0, 11, 2, O.ptr(d), 1.0, 3.0, 0, O.ptr(out := np.zeros_like(d)))
    assert out[5, 0] == pytest.approx(10 * w[3] / w.sum(), rel=1e-12)


# ---------------------------------------------------------------- slab / regions / molecules
# mrmd/weighting_function/Slab.test.cpp:23-80, CheckRegion.test.cpp:23-42
def _slab_eval(s, x):
    out = np.zeros(5)
    O.lib().or_slab_eval(O.C.byref(s), x, O.ptr(out))
    return out


def test_slab_regions_and_values():
    s = O.make_slab([2, 3, 4], 2.0, 2.0, 2)
    at = O.pred1(kind=4, a=2.0, b=1.0)
    hy = O.pred1(kind=5, a=2.0, b=1.0, c=2.0)
    cg = O.pred1(kind=6, a=2.0, b=1.0, c=2.0)
    # region predicates evaluated through the Langevin/thermodynamic-force predicate path
    assert _slab_eval(s, 2.0)[1] == 1.0
    assert 0 < _slab_eval(s, 3.1)[1] < 1
    assert _slab_eval(s, 5.2)[1] == 0.0
    assert _slab_eval(s, 2.9)[1] == 1.0
    assert _slab_eval(s, 5.1)[1] == 0.0
    w = [_slab_eval(s, 2.0 + 0.1 * i)[1] for i in range(60)]
    assert all(w[i + 1] <= w[i] for i in range(59))
    del at, hy, cg


def test_slab_gradient_finite_difference():
    s = O.make_slab([10, 0, 0], 4.0, 2.0, 1)
    for x in np.linspace(12.05, 13.95, 25):
        h = 1e-6
        lp = _slab_eval(s, x + h)[1]
        lm = _slab_eval(s, x - h)[1]
        assert _slab_eval(s, x)[2] == pytest.approx((lp - lm) / (2 * h), rel=1e-6, abs=1e-9)
        assert _slab_eval(s, 20 - x)[2] == pytest.approx(-_slab_eval(s, x)[2], rel=1e-12)


# mrmd/action/UpdateMolecules.test.cpp:43-60, ContributeMoleculeForceToAtoms.test.cpp:27-51,
# mrmd/test/DiamondFixture.hpp (2 molecules x 2 atoms, relMass 0.25 / 0.75)
def _diamond():
    pos = np.array([[1.0, 0, 0], [0, 1, 0], [-1, 0, 0], [0, -1, 0]])
    atoms = O.Atoms.from_arrays(pos, relmass=np.array([0.25, 0.75, 0.25, 0.75]))
    mols = O.Molecules(2)
    mols.set(O.M_OFFSET, [0, 2])
    mols.set(O.M_NUMATOMS, [2, 2])
    mols.set_counts(2, 0)
    return mols, atoms


def test_update_molecules_com():
    mols, atoms = _diamond()
    O.lib().or_update_molecules(mols.h, atoms.h, None)
    np.testing.assert_allclose(mols.get(O.M_POS), [[0.25, 0.75, 0], [-0.25, -0.75, 0]])


def test_contribute_molecule_force():
    mols, atoms = _diamond()
    mols.set(O.M_FORCE, [[5, 6, 7], [0, 0, 0]])
    O.lib().or_contribute_molecule_force(mols.h, atoms.h)
    F = atoms.get(O.FORCE)
    np.testing.assert_allclose(F[0], [1.25, 1.5, 1.75])
    np.testing.assert_allclose(F[1], [3.75, 4.5, 5.25])


# mrmd/communication/MultiResPeriodicGhostExchange.test.cpp:25-62, GridFixture.hpp
def _grid_fixture(apm=2):
    box = O.make_box([0, 0, 0], [3, 3, 3], 0.7)
    mp, ap = [], []
    for x in range(3):
        for y in range(3):
            for z in range(3):
                mp.append((x + 0.5, y + 0.5, z + 0.5))
                for i in range(apm):
                    ap.append((x + 0.5 + 0.1 * i, y + 0.5 + 0.2 * i, z + 0.5 + 0.3 * i))
    mols = O.Molecules(27)
    mols.set(O.M_POS, mp)
    mols.set(O.M_OFFSET, np.arange(27) * apm)
    mols.set(O.M_NUMATOMS, np.full(27, apm))
    mols.set_counts(27, 0)
    atoms = O.Atoms.from_arrays(np.array(ap))
    return box, mols, atoms


@pytest.mark.parametrize("axis", [0, 1, 2])
def test_multires_ghost_axis(axis):
    box, mols, atoms = _grid_fixture()
    g = O.Ghost()
    g.reset(atoms)
    O.lib().or_multires_ghost_axis(g.h, mols.h, atoms.h, O.C.byref(box), axis)
    assert mols.nghost == 2 * 3 * 3
    assert atoms.nghost == 3 * 3 * 2 * 2


def test_multires_ghost_xyz():
    box, mols, atoms = _grid_fixture()
    g = O.Ghost()
    O.lib().or_multires_ghost_xyz(g.h, mols.h, atoms.h, O.C.byref(box))
    assert mols.nghost == 5 * 5 * 5 - 3 * 3 * 3
    assert atoms.nghost == (5 * 5 * 5 - 3 * 3 * 3) * 2


@pytest.mark.parametrize("apm", [1, 2])
def test_multires_real_atoms_exchange(apm):
    _, mols, atoms = _grid_fixture(apm)
    box = O.make_box([1, 1, 1], [2, 2, 2], 0.1)
    O.lib().or_multires_exchange_real(O.C.byref(box), mols.h, atoms.h)
    mp = mols.get(O.M_POS)[:27]
    assert np.all(mp >= 1) and np.all(mp < 2)
    ap = atoms.get(O.POS)[: 27 * apm]
    assert np.all(ap >= 1) and np.all(ap < 2)


# ---------------------------------------------------------------- spherical weighting
# mrmd/weighting_function/Spherical.test.cpp:22-85 (centre (2, 3, 4), atomistic radius 2,
# hybrid shell 2, exponent 2)
def _sph(x, y, z, center=(2.0, 3.0, 4.0), rat=2.0, hy=2.0, n=2):
    out = np.zeros(4)
    c = np.asarray(center, dtype=np.float64)
    O.lib().or_spherical_eval(O.ptr(c), rat, hy, n, x, y, z, O.ptr(out))
    return out


def test_spherical_reference_cases():
    # monotonous along the diagonal from the centre
    w = [_sph(2.0 + 0.1 * i, 3.0 + 0.1 * i, 4.0 + 0.1 * i)[0] for i in range(60)]
    assert all(w[i + 1] <= w[i] for i in range(59))
    # boundary values
    assert _sph(2.1, 3.1, 4.1)[0] == 1.0
    assert _sph(6.1, 7.1, 8.1)[0] == 0.0
    # derivative as the reference test evaluates it (x varied at y = z = 0)
    eps = 1e-10
    for i in range(60):
        x = 2.0 + 0.1 * i
        l0, l1 = _sph(x, 0.0, 0.0), _sph(x + eps, 0.0, 0.0)
        assert (l1[0] - l0[0]) / eps == pytest.approx(0.5 * (l0[1] + l1[1]), abs=1e-6)


def test_spherical_closed_form_and_gradient():
    """inside the shell lambda = cos^n(pi/(2 d_HY) (r - r_AT)); the gradient is the radial
    derivative times the unit vector (central differences along every axis)"""
    g = np.random.default_rng(3)
    c = np.array([2.0, 3.0, 4.0])
    for _ in range(200):
        u = g.normal(size=3)
        u /= np.linalg.norm(u)
        r = g.uniform(2.0, 4.0)
        p = c + r * u
        out = _sph(*p, n=3)
        assert out[0] == pytest.approx(math.cos(math.pi / 4.0 * (r - 2.0)) ** 3, rel=1e-12, abs=1e-15)
        h = 1e-6
        for d in range(3):
            e = np.zeros(3)
            e[d] = h
            fd = (_sph(*(p + e), n=3)[0] - _sph(*(p - e), n=3)[0]) / (2 * h)
            assert out[1 + d] == pytest.approx(fd, rel=1e-6, abs=1e-8)
    # a different exponent changes lambda as cos^n
    p = c + np.array([3.0, 0.0, 0.0])
    assert _sph(*p, n=1)[0] ** 2 == pytest.approx(_sph(*p, n=2)[0], rel=1e-14)


# ---------------------------------------------------------------- Berendsen thermostat
# mrmd/action/BerendsenThermostat.test.cpp:24-45 on the single-atom fixture (T = 41.5)
@pytest.mark.parametrize("gamma,expected", [(0.0, 41.5), (1.0, 3.8)])
def test_berendsen_reference_cases(gamma, expected):
    a = _single_atom()
    t = O.lib().or_kinetic_energy(a.h) * (2.0 / 3.0)
    assert t == pytest.approx(41.5, rel=1e-15)
    assert O.lib().or_berendsen_apply(a.h, t, 3.8, gamma) == 0
    assert O.lib().or_kinetic_energy(a.h) * (2.0 / 3.0) == pytest.approx(expected, rel=1e-12)


def test_berendsen_scaling_law():
    """v scales by sqrt(1 + gamma (T0/T - 1)), so T moves a fraction gamma towards T0; no
    change for T <= 0; a non-positive target is rejected"""
    g = np.random.default_rng(4)
    v = g.normal(size=(1000, 3))
    a = O.Atoms.from_arrays(g.uniform(0, 10, (1000, 3)), vel=v)
    O.lib().or_berendsen_apply(a.h, 2.0, 1.0, 0.25)
    np.testing.assert_allclose(a.get(O.VEL), v * math.sqrt(1.0 + 0.25 * (1.0 / 2.0 - 1.0)), rtol=1e-15)
    before = a.get(O.VEL)
    assert O.lib().or_berendsen_apply(a.h, 0.0, 1.0, 0.5) == 0
    np.testing.assert_array_equal(a.get(O.VEL), before)
    assert O.lib().or_berendsen_apply(a.h, 1.0, 0.0, 0.5) == -1
