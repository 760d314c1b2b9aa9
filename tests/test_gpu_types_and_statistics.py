"""Multi-type LennardJones with the reference's own type index, and the reference's
statistical integration tests, on the GPU.

* LennardJones.cpp:52 keeps numTypes_ = 1, so the reference's pair index is t_i + t_j
  (type stride 1); LJ_IdealGas and the correct form use t_i * n + t_j (stride n).  Both run
  on the full-row kernel (stride 1 is always symmetric in i and j) and on the half-list
  kernel (an asymmetric pair predicate, or an asymmetric stride-n table) against the oracle.
* tests/LangevinThermostat/LangevinThermostat.cpp:82-136: 1e5 free atoms, zeta = 1e5,
  dt = 1e-3, 21 steps relax to T = 1.12 +- 0.01, globally and with the thermostat applied
  through a symmetric-slab predicate.
* tests/NVT/NVT.cpp:118-193: 4,096 atoms at density 0.8442 from uniform-random positions,
  capped unshifted LJ, velocity Verlet with the Berendsen thermostat driven by the moving
  average of T, 2,001 steps per temperature: the moving-average pressure 2 (E_kin - W) /
  (3 V) lies within 0.3 of the reference's fit p(T) = -0.89528939 T^2 + 7.48553466 T -
  4.00636731 and T within 0.1 of the target, for T = 0.8 ... 1.6.
"""
import numpy as np
import pytest

import pyoracle as O
from mrmd_amd.lattice import fcc, rng, uniform

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


def _two_type_system(M, ctx):
    pos, L = fcc(7, 0.8442)
    g = np.random.default_rng(12)
    pos = np.mod(pos + g.uniform(-0.12, 0.12, pos.shape), L)
    types = (g.random(pos.shape[0]) < 0.35).astype(np.int64)
    ga = M.Atoms.from_arrays(ctx, pos, type=types)
    ca = O.Atoms.from_arrays(pos, type=types)
    gl = M.GhostLayer(ctx)
    gl.create_ghost_atoms(ga, M.make_box([0, 0, 0], [L] * 3, 2.8))
    og = O.Ghost()
    og.create_xyz(ca, O.make_box([0, 0, 0], [L] * 3, 2.8))
    vl = M.HalfVerletList(ctx)
    vl.build(ga, 0, ga.nlocal, 2.8, 1.0, [-2.8] * 3, [L + 2.8] * 3)
    ov = O.Verlet()
    ov.build(ca, ca.nlocal + ca.nghost, 0, ca.nlocal, 2.8, 1.0, [-2.8] * 3, [L + 2.8] * 3)
    return L, ga, ca, gl, og, vl, ov


@pytest.mark.parametrize("stride,table,pred,kernel", [
    (1, "symmetric", False, "lj_force"),
    (1, "symmetric", True, "lj_force_half"),
    (2, "symmetric", False, "lj_force"),
    (2, "asymmetric", False, "lj_force_half"),
], ids=["reference-index-full-rows", "reference-index-half-list", "type-major-full-rows",
        "type-major-asymmetric-half-list"])
def test_two_type_lennard_jones_type_index(M, ctx, stride, table, pred, kernel):
    L, ga, ca, gl, og, vl, ov = _two_type_system(M, ctx)
    cap, rc = [0.8] * 4, [2.5] * 4
    if table == "symmetric":
        sig, eps = [1.0, 0.93, 0.93, 0.86], [1.0, 0.75, 0.75, 0.5]
    else:
        sig, eps = [1.0, 0.93, 0.89, 0.86], [1.0, 0.75, 0.65, 0.5]
    lj = M.LennardJones.vector(ctx, cap, rc, sig, eps, 2, True) if stride == 1 else \
        M.LennardJones(ctx, rc, sig, eps, capping_distance=cap, num_types=2, shifted=True, type_stride=2)
    olj = O.LJ(cap, rc, sig, eps, 2, shifted=True, type_stride=stride)
    np.testing.assert_array_equal(lj.table(), olj.table())
    gp = cp = None
    if pred:
        region = dict(kind=M.PRED_OPEN_RANGE, axis=0, a=0.2 * L, b=0.7 * L)
        other = dict(kind=M.PRED_LESS, axis=2, a=0.6 * L)
        gp = M.pred2(M.PAIR_AND, M.pred1(**region), M.pred1(**other))
        cp = O.pred2(1, O.pred1(**region), O.pred1(**other))
    ctx.set_profiling(True)
    ctx.reset_stats()
    ga.set_force(0.0)
    lj.apply(ga, vl, gp)
    gl.contribute_back_ghost_to_real(ga)
    assert ctx.kernel_stats(kernel)[0] == 1
    ctx.set_profiling(False)
    cE, cW = olj.apply(ca, ov, cp)
    og.fold(ca)
    E, W = lj.energy_virial()
    assert E == pytest.approx(cE, rel=1e-12) and W == pytest.approx(cW, rel=1e-12)
    Fg, Fc = ga.download(M.FORCE, ga.nlocal), ca.get(O.FORCE, ca.nlocal)
    assert np.abs(Fg - Fc).max() <= 1e-12 * np.abs(Fc).max()
    if stride == 1 and not pred:
        # the index matters: the same table read type-major gives other forces
        olj2 = O.LJ(cap, rc, sig, eps, 2, shifted=True, type_stride=2)
        c2 = O.Atoms.from_arrays(ca.get(O.POS), type=ca.get(O.TYPE))
        c2.set_counts(ca.nlocal, ca.nghost)
        olj2.apply(c2, ov)
        assert np.abs(c2.get(O.FORCE, ca.nlocal) - Fc).max() > 1e-3


@pytest.mark.parametrize("local", [False, True], ids=["global", "symmetric-slab"])
def test_langevin_thermostat_relaxes_to_target(M, ctx, local):
    n, Lx, dt, T0 = 100000, 10.0, 1e-3, 1.12
    pos = uniform(n, [0, 0, 0], [Lx] * 3, 1234)
    vel = (rng(4321).random((n, 3)) - 0.5) * 10.0
    a = M.Atoms.from_arrays(ctx, pos, vel=vel)
    th = M.VelocityVerletLangevinThermostat(1e5, T0, seed=1234)
    p = M.pred1(kind=M.PRED_SYMMETRIC_SLAB, axis=0, a=0.5 * Lx, b=0.0, c=5.0) if local else None
    for _ in range(21):
        th.pre_force_integrate(a, dt, p)
        a.set_force(0.0)
        th.post_force_integrate(a, dt)
    T = (2.0 / 3.0) * a.kinetic_energy() / n
    assert abs(T - T0) < 0.01, T


class _EMA:
    """util::ExponentialMovingAverage (ExponentialMovingAverage.hpp:23-52)"""

    def __init__(self, alpha):
        self.alpha, self.val, self.first = alpha, 0.0, True

    def append(self, v):
        if self.first:
            self.first, self.val = False, v
        else:
            self.val = v * self.alpha + self.val * (1.0 - self.alpha)


def test_nvt_equation_of_state_berendsen(M, ctx):
    Lx, rho, rc, skin, dt = 16.92926877476863, 0.8442, 2.5, 0.3, 0.005
    rn = rc + skin
    volume = Lx ** 3
    n = int(rho * volume)
    box = M.make_box([0, 0, 0], [Lx] * 3, rn)
    for k, target in enumerate((0.8, 1.0, 1.2, 1.4, 1.6)):
        pos = uniform(n, [0, 0, 0], [Lx] * 3, 100 + k)
        vel = (rng(200 + k).random((n, 3)) - 0.5) * 1.0
        a = M.Atoms.from_arrays(ctx, pos, vel=vel)
        gl = M.GhostLayer(ctx)
        vl = M.HalfVerletList(ctx)
        cells = None
        lj = M.LennardJones(ctx, rc, 1.0, 1.0, capping_distance=0.7)  # 4-argument form: unshifted
        max_disp = np.finfo(np.float64).max
        p_ema, t_ema = _EMA(0.01), _EMA(0.01)
        for step in range(2001):
            max_disp += M.VelocityVerlet.pre_force_integrate(a, dt)
            if max_disp >= 0.5 * skin:
                max_disp = 0.0
                gl.exchange_real_atoms(a, box)
                if cells is None:
                    cells = M.LinkedCellList(ctx, a, 0, a.nlocal, [rn] * 3, [0, 0, 0], [Lx] * 3)
                else:
                    cells.build(a, 0, a.nlocal, [rn] * 3, [0, 0, 0], [Lx] * 3)
                a.permute(cells)
                gl.create_ghost_atoms(a, box)
                vl.build(a, 0, a.nlocal, rn, 1.0, [-rn] * 3, [Lx + rn] * 3, 60)
            else:
                gl.update_ghost_atoms(a, box)
            a.set_force(0.0)
            lj.apply(a, vl)
            # outputInterval = -1: step % -1 == 0 on every step; the averages restart on
            # every step before 201 (NVT.cpp:165-177)
            if step < 201:
                p_ema, t_ema = _EMA(0.1), _EMA(0.1)
            _, W = lj.energy_virial()
            ek = a.kinetic_energy()
            p_ema.append(2.0 * (ek - W) / (3.0 * volume))
            t_ema.append((2.0 / 3.0) * ek / n)
            M.BerendsenThermostat.apply(a, t_ema.val, target, 1.0)
            gl.contribute_back_ghost_to_real(a)
            M.VelocityVerlet.post_force_integrate(a, dt)
        fit = -0.89528939 * target ** 2 + 7.48553466 * target - 4.00636731
        assert abs(p_ema.val - fit) < 0.3, (target, p_ema.val, fit)
        assert abs(t_ema.val - target) < 0.1, (target, t_ema.val)
