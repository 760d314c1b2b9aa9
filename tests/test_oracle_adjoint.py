"""Pins the CPU adjoint of the LJ force map (oracle.cpp: or_lj_adjoint) by central finite
differences (h = 1e-6) of L = sum_i g_i . F_i in positions and in the pair parameters.
The FD derivative is formed from per-atom force differences, so its round-off is far
below the 1e-6 agreement demanded."""
import numpy as np
import pytest

import pyoracle as O
from mrmd_amd.lattice import fcc


def _lb(eA, eB, sA, sB):
    eAB = np.sqrt(eA * eB)
    sAB = 0.5 * (sA + sB)
    return np.array([eA, eAB, eAB, eB]), np.array([sA, sAB, sAB, sB])


def _system(seed=3, cells=4, rho=0.8):
    pos, L = fcc(cells, rho)
    rng = np.random.default_rng(seed)
    pos = np.mod(pos + rng.uniform(-0.12, 0.12, pos.shape), L)
    types = (rng.random(pos.shape[0]) < 0.2).astype(np.int64)
    return pos, types, L


def _forces(pos, types, L, eps, sig, vl=None, keep=None):
    a = O.Atoms.from_arrays(pos, type=types)
    box = O.make_box([0, 0, 0], [L] * 3, 2.8)
    g = O.Ghost()
    g.create_xyz(a, box)
    if vl is None:
        vl = O.Verlet()
        vl.build(a, a.nlocal + a.nghost, 0, a.nlocal, 2.8, 1.0, [-2.8] * 3, [L + 2.8] * 3)
    lj = O.LJ([0.0] * 4, [2.5] * 4, sig, eps, 2, shifted=True, type_stride=2)
    lj.apply(a, vl)
    g.fold(a)
    return a.get(O.FORCE, a.nlocal), a, g, vl, lj


@pytest.fixture(autouse=True, scope="module")
def _serial():
    O.lib().or_set_threads(1)
    yield


def test_adjoint_matches_finite_differences():
    pos, types, L = _system()
    n = pos.shape[0]
    eps, sig = _lb(1.0, 0.5, 1.0, 0.88)
    F0, a, g, vl, lj = _forces(pos, types, L, eps, sig)
    up = np.random.default_rng(9).standard_normal((n, 3))
    dx = np.zeros((n, 3))
    de = np.zeros(4)
    ds = np.zeros(4)
    O.lib().or_lj_adjoint(lj.h, a.h, vl.h, g.h, O.ptr(up), O.ptr(dx), O.ptr(de), O.ptr(ds))
    h = 1e-6
    # positions: a few random coordinates (ghost images follow through re-creation)
    rng = np.random.default_rng(1)
    for k in rng.choice(n, 12, replace=False):
        for d in range(3):
            pp, pm = pos.copy(), pos.copy()
            pp[k, d] += h
            pm[k, d] -= h
            Fp = _forces(pp, types, L, eps, sig)[0]
            Fm = _forces(pm, types, L, eps, sig)[0]
            fd = np.sum(up * (Fp - Fm)) / (2 * h)
            assert dx[k, d] == pytest.approx(fd, rel=1e-6, abs=1e-7)
    # symmetric pair parameters: eps_AA (0), eps_AB (1 = 2), eps_BB (3); same for sigma
    for which, grad in (("eps", de), ("sig", ds)):
        for t, entries in ((0, [0]), (1, [1, 2]), (3, [3])):
            ep, em = eps.copy(), eps.copy()
            sp, sm = sig.copy(), sig.copy()
            for e in entries:
                if which == "eps":
                    ep[e] += h
                    em[e] -= h
                else:
                    sp[e] += h
                    sm[e] -= h
            Fp = _forces(pos, types, L, ep, sp, vl=None)[0]
            Fm = _forces(pos, types, L, em, sm, vl=None)[0]
            fd = np.sum(up * (Fp - Fm)) / (2 * h)
            assert grad[t] == pytest.approx(fd, rel=1e-6), (which, t)
    assert de[1] == de[2] and ds[1] == ds[2]
