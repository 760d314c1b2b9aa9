"""Pins of the oracle's AdResS LJ / ideal-gas terms that the reference's own 4-atom test
leaves open (its molecules have grad lambda = 0): the drift force -V_ij grad(lambda_mod),
the compensation-energy samples and counts, updateMeanCompensationEnergy, and the mean
compensation force +meanComp grad(lambda_mod) (action/LJ_IdealGas.cpp:21-50, 52-215).

* Gradient identity.  With the mean compensation energy zero, LJ_IdealGas::run computes the
  forces of U = sum over list pairs of w_ab E_ij, w_ab = (lambda_mod,a + lambda_mod,b) / 2:
  the atom force w F_ij plus, after ContributeMoleculeForceToAtoms (relMass-weighted) and
  the ghost fold, the drift -1/2 sum E_ij grad(lambda_mod) of each molecule.  So the atom
  forces must equal -grad U by central differences of the returned energy.  Exponent nu = 2
  makes lambda_mod = cos^4 differ from lambda = cos^2, so using grad(lambda) or lambda where
  lambda_mod belongs fails; a dropped or sign-flipped drift term fails by O(0.1-1).
* Compensation.  Three molecules on an equilateral triangle of side 2^(1/6) sigma with an
  unshifted potential: every pair energy is -epsilon, so every V_ij = -epsilon / 2 and each
  hybrid molecule's bin collects -epsilon per sampling run - literal numbers.  The running
  mean after two updates is (10 m0 + e) / 11 twice over.
* Regions and types.  Five molecules in the atomistic, hybrid and coarse-grained regions,
  two types, lambda != lambda_mod, one diatomic: samples go to the bin of lambda (not
  lambda_mod) and the column of each atom's type, counts add one per atom of a hybrid local
  molecule, pairs of two coarse-grained molecules are skipped, the drift acts only when one
  of the pair is hybrid.
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


def _ljig_forces(atoms, mols, vl, lj, slab, ghost=None, box=None):
    """energy and local atom forces of one LJ_IdealGas evaluation with the mean
    compensation energy zero: ghost refresh, UpdateMolecules with the slab, run,
    ContributeMoleculeForceToAtoms, ghost fold"""
    if ghost is not None:
        ghost.update(atoms, box)
    O.lib().or_update_molecules(mols.h, atoms.h, O.C.byref(slab))
    O.lib().or_atoms_set_force(atoms.h, 0.0)
    O.lib().or_mol_set_force(mols.h, 0.0)
    lj.set_mean(np.zeros((200, lj.ntypes)))
    e = lj.run(mols, vl, atoms)
    O.lib().or_contribute_molecule_force(mols.h, atoms.h)
    if ghost is not None:
        ghost.fold(atoms)
    return e, atoms.get(O.FORCE, atoms.nlocal)


def _fd_check(atoms, mols, vl, lj, slab, rows, ghost=None, box=None, h=1e-6):
    e0, F = _ljig_forces(atoms, mols, vl, lj, slab, ghost, box)
    n = atoms.nlocal
    pos = atoms.get(O.POS, n)
    fd = np.zeros((len(rows), 3))
    for k, i in enumerate(rows):
        for d in range(3):
            for sgn in (1.0, -1.0):
                p = pos.copy()
                p[i, d] += sgn * h
                atoms.set(O.POS, p)
                e, _ = _ljig_forces(atoms, mols, vl, lj, slab, ghost, box)
                fd[k, d] -= sgn * e / (2.0 * h)
    atoms.set(O.POS, pos)
    return e0, F[rows], fd


def _assert_fd(F, fd):
    """central differences with h = 1e-6 carry ~1e-8 of rounding (energies of order 1e3 over
    2h) and ~1e-10 of truncation here; the measured gap is below 2e-7 of (1 + |F|)"""
    err = np.abs(F - fd) / (1.0 + np.abs(F))
    assert err.max() <= 1e-6, err.max()


def _jittered_grid(shape, spacing, jitter, seed):
    g = np.random.default_rng(seed)
    idx = np.stack(np.meshgrid(*[np.arange(s) for s in shape], indexing="ij"), -1).reshape(-1, 3)
    return (idx + 0.5) * spacing + g.uniform(-jitter, jitter, idx.shape)


def _hybrid_rows(com, centre, half_at, hy, want=36):
    ax = np.abs(com[:, 0] - centre)
    hyb = np.where((ax > half_at + 1e-3) & (ax < half_at + hy - 1e-3))[0]
    at = np.where(ax < half_at)[0]
    cg = np.where(ax > half_at + hy)[0]
    rows = list(hyb[:: max(1, len(hyb) // want)][:want]) + list(at[:3]) + list(cg[:3])
    assert len(hyb) >= 12
    return rows


@pytest.mark.parametrize("nu", [1, 2])
def test_gradient_identity_diatomic_open(nu):
    """open box (no ghosts), 400 diatomic molecules (relMass 1/2 each, bond 0.4 sigma),
    slab centred at x = 12: atomistic |x - 12| < 3, hybrid width 5"""
    com = _jittered_grid((16, 5, 5), 1.5, 0.1, 3)
    g = np.random.default_rng(4)
    u = g.normal(size=com.shape)
    u /= np.linalg.norm(u, axis=1, keepdims=True)
    pos = np.empty((2 * com.shape[0], 3))
    pos[0::2] = com + 0.2 * u
    pos[1::2] = com - 0.2 * u
    n_mol = com.shape[0]
    atoms = O.Atoms.from_arrays(pos, relmass=np.full(pos.shape[0], 0.5))
    mols = O.Molecules(n_mol)
    mols.set(O.M_OFFSET, np.arange(n_mol) * 2)
    mols.set(O.M_NUMATOMS, np.full(n_mol, 2))
    mols.set_counts(n_mol, 0)
    slab = O.make_slab([12.0, 0, 0], 6.0, 5.0, nu)
    O.lib().or_update_molecules(mols.h, atoms.h, O.C.byref(slab))
    vl = O.Verlet()
    rn = 2.5 + 0.4 + 0.3  # atom pairs within rc have centres within rc + bond length
    vl.build(mols, n_mol, 0, n_mol, rn, 1.0, [-1.0, -1.0, -1.0], [25.0, 8.5, 8.5], molecules=True)
    lj = O.LJIdealGas([0.0], [2.5], [1.0], [1.0], 1, shifted=True)
    mol_rows = _hybrid_rows(mols.get(O.M_POS, n_mol), 12.0, 3.0, 5.0, want=20)
    rows = [2 * m for m in mol_rows] + [2 * m + 1 for m in mol_rows[:10]]
    e0, F, fd = _fd_check(atoms, mols, vl, lj, slab, rows)
    _assert_fd(F, fd)
    # the drift is a sizeable part of the force that is checked
    lam = mols.get(O.M_GRAD, n_mol)[mol_rows, 0]
    assert np.abs(lam).max() > 0.1


def test_gradient_identity_monatomic_periodic_slab_across_boundary():
    """periodic 24 x 7.2 x 7.2 box, 1,440 monatomic molecules, multi-resolution ghosts, slab
    centred at x = 2 so that hybrid ghost images face coarse-grained reals; forces after the
    ghost fold against -dU/dx of each real atom (its ghost images move with it)"""
    L = [24.0, 7.2, 7.2]
    pos = np.mod(_jittered_grid((20, 6, 6), 1.2, 0.15, 8), L)
    atoms = O.Atoms.from_arrays(pos)
    mols = O.Molecules.for_each_atom(atoms)
    rn = 2.8
    box = O.make_box([0, 0, 0], L, rn)
    ghost = O.Ghost()
    O.lib().or_update_molecules(mols.h, atoms.h, None)
    O.lib().or_multires_ghost_xyz(ghost.h, mols.h, atoms.h, O.C.byref(box))
    O.lib().or_update_molecules(mols.h, atoms.h, None)
    nm = mols.nlocal + mols.nghost
    vl = O.Verlet()
    vl.build(mols, nm, 0, mols.nlocal, rn, 1.0, [-rn] * 3, [L[0] + rn, L[1] + rn, L[2] + rn], molecules=True)
    slab = O.make_slab([2.0, 0, 0], 6.0, 5.0, 2)
    lj = O.LJIdealGas([0.0], [2.5], [1.0], [1.0], 1, shifted=True)
    O.lib().or_update_molecules(mols.h, atoms.h, O.C.byref(slab))
    lam = mols.get(O.M_MODLAMBDA, nm)
    assert (lam[mols.nlocal:] > 0).any()
    x = pos[:, 0]
    dx = np.minimum(np.abs(x - 2.0), L[0] - np.abs(x - 2.0))
    hyb = np.where((dx > 3.001) & (dx < 7.999))[0]
    rows = list(hyb[:: max(1, len(hyb) // 30)][:30])
    # rows whose images sit across the boundary (x > L - 8 or x < 8 from the centre at 2)
    assert any(x[r] > 16.0 for r in rows) and any(x[r] < 8.0 for r in rows)
    e0, F, fd = _fd_check(atoms, mols, vl, lj, slab, rows, ghost, box)
    _assert_fd(F, fd)


def test_gradient_identity_fails_for_a_wrong_drift_sign():
    """the identity separates the right drift term from its negation: adding twice the
    drift that run() applies (i.e. flipping its sign) moves the forces far outside the
    tolerance the identity tests hold"""
    com = _jittered_grid((14, 5, 5), 1.2, 0.15, 5)
    atoms = O.Atoms.from_arrays(com)
    mols = O.Molecules.for_each_atom(atoms)
    n = com.shape[0]
    slab = O.make_slab([8.4, 0, 0], 4.0, 4.0, 2)
    O.lib().or_update_molecules(mols.h, atoms.h, O.C.byref(slab))
    vl = O.Verlet()
    vl.build(mols, n, 0, n, 2.8, 1.0, [-1.0] * 3, [18.0, 7.0, 7.0], molecules=True)
    lj = O.LJIdealGas([0.0], [2.5], [1.0], [1.0], 1, shifted=True)
    rows = _hybrid_rows(mols.get(O.M_POS, n), 8.4, 2.0, 4.0, want=12)
    _, F, fd = _fd_check(atoms, mols, vl, lj, slab, rows)
    # the drift alone: molecule forces of run() minus the LJ part (atom forces before
    # ContributeMoleculeForceToAtoms)
    O.lib().or_atoms_set_force(atoms.h, 0.0)
    O.lib().or_mol_set_force(mols.h, 0.0)
    lj.set_mean(np.zeros((200, 1)))
    lj.run(mols, vl, atoms)
    drift = mols.get(O.M_FORCE, n)[rows]
    _assert_fd(F, fd)
    assert (np.abs((F - 2.0 * drift) - fd) / (1.0 + np.abs(F))).max() > 1e-2


def _molecules(lams, modlams, grads, nats=None):
    k = len(lams)
    nats = [1] * k if nats is None else nats
    mols = O.Molecules(k)
    mols.set(O.M_OFFSET, np.concatenate([[0], np.cumsum(nats)[:-1]]))
    mols.set(O.M_NUMATOMS, nats)
    mols.set(O.M_LAMBDA, lams)
    mols.set(O.M_MODLAMBDA, modlams)
    mols.set(O.M_GRAD, grads)
    mols.set_counts(k, 0)
    return mols


def test_compensation_energy_triangle_literal():
    """three hybrid monatomic molecules on an equilateral triangle of side 2^(1/6), LJ
    epsilon = sigma = 1 unshifted: every E_ij = -1, V_ij = -1/2.  Sampling every run,
    update every second run (set_intervals(1, 2)):
      run 0: forces with mean 0; samples -1 per bin, count 1; update -> mean = -1/11
      run 1: forces with mean -1/11; samples -1, count 1 (no update)
      run 2: update with comp -2, count 2 -> mean = (10 (-1/11) + (-1)) / 11 = -21/121"""
    s = 2.0 ** (1.0 / 6.0)
    pos = np.array([[0.0, 0.0, 0.0], [s, 0.0, 0.0], [0.5 * s, 0.5 * math.sqrt(3.0) * s, 0.0]])
    lam = [0.3013, 0.5527, 0.8061]          # bins 60, 110, 161 of 200 over [0, 1)
    mlam = [0.4, 0.6, 0.7]                  # lambda_mod sets the weights and the region
    grad = [[-0.5, 0.25, 0.0], [0.75, 0.0, -0.125], [0.0, -1.5, 0.5]]
    mols = _molecules(lam, mlam, grad)
    mols.set(O.M_POS, pos)
    atoms = O.Atoms.from_arrays(pos)
    vl = O.Verlet()
    vl.build(mols, 3, 0, 3, 2.0, 1.0, [-2.0] * 3, [3.0] * 3, molecules=True)
    assert vl.total() == 3
    lj = O.LJIdealGas([0.0], [2.5], [1.0], [1.0], 1, shifted=False)
    lj.set_intervals(1, 2)
    bins = [60, 110, 161]
    G = np.array(grad)

    def run():
        O.lib().or_atoms_set_force(atoms.h, 0.0)
        O.lib().or_mol_set_force(mols.h, 0.0)
        e = lj.run(mols, vl, atoms)
        return e, mols.get(O.M_FORCE, 3)

    # run 0: E = sum w E_ij, w_ab = (ml_a + ml_b) / 2; drift per molecule = -sum V grad =
    # +1 * grad (two pairs of -(-1/2))
    e, Fm = run()
    w = [(mlam[0] + mlam[1]) / 2, (mlam[0] + mlam[2]) / 2, (mlam[1] + mlam[2]) / 2]
    assert e == pytest.approx(-sum(w), abs=1e-14)
    np.testing.assert_allclose(Fm, G, atol=1e-14)
    # the LJ forces vanish at the minimum
    assert np.abs(atoms.get(O.FORCE, 3)).max() < 1e-12
    m1 = np.zeros((200, 1))
    m1[bins, 0] = -1.0 / 11.0
    np.testing.assert_allclose(lj.hist(2), m1, atol=1e-15)
    assert not lj.hist(0).any() and not lj.hist(1).any()
    # run 1: mean compensation force m(bin(lambda)) grad added
    e, Fm = run()
    np.testing.assert_allclose(Fm, G * (1.0 - 1.0 / 11.0), atol=1e-14)
    c = np.zeros((200, 1))
    c[bins, 0] = -1.0
    np.testing.assert_allclose(lj.hist(0), c, atol=1e-14)
    cnt = np.zeros((200, 1))
    cnt[bins, 0] = 1.0
    np.testing.assert_array_equal(lj.hist(1), cnt)
    # run 2: update from two samples
    run()
    m2 = np.zeros((200, 1))
    m2[bins, 0] = -21.0 / 121.0
    np.testing.assert_allclose(lj.hist(2), m2, atol=1e-15)
    assert not lj.hist(0).any() and not lj.hist(1).any()


def _lj_pair(r2, eps, sig, rc):
    """closed-form LJ force factor F / r and energy, shifted so that E(rc) = 0"""
    def e(x2):
        s6 = (sig * sig / x2) ** 3
        return 4.0 * eps * (s6 * s6 - s6)
    s6 = (sig * sig / r2) ** 3
    return 24.0 * eps * (2.0 * s6 * s6 - s6) / r2, e(r2) - e(rc * rc)


def test_compensation_regions_types_and_counts():
    """five molecules: A atomistic, H1 / H2 hybrid (H2 diatomic), C1 / C2 coarse-grained;
    two types with a 2 x 2 table, shifted.  Expected values from the closed-form LJ pair
    terms: samples V = E / 2 go to (bin(lambda), type of the sampled side's atom) for each
    hybrid side, counts add one per atom of a hybrid molecule, the C1-C2 pair is skipped,
    the A-C pairs have no drift (neither side hybrid)."""
    eps = np.array([[1.0, 0.7], [0.7, 0.5]])
    sig = np.array([[1.0, 0.95], [0.95, 0.9]])
    rc = 2.5
    #          A              H1              H2 (2 atoms)                     C1             C2
    apos = np.array([[0.0, 0, 0], [1.1, 0.2, 0], [0.4, 1.15, 0.1], [0.9, 1.3, 1.05], [0.2, 0.3, 1.2],
                     [1.3, 0.1, 1.3]])
    atype = np.array([0, 1, 0, 1, 1, 0])
    nats = [1, 1, 2, 1, 1]
    off = [0, 1, 2, 4, 5]
    lam = [1.0, 0.4231, 0.6517, 0.0, 0.0]
    mlam = [1.0, 0.35, 0.62, 0.0, 0.0]
    grad = [[0, 0, 0], [0.8, -0.2, 0.1], [-0.6, 0.3, 0.0], [0, 0, 0], [0, 0, 0]]
    mols = _molecules(lam, mlam, grad, nats)
    com = np.array([apos[o:o + k].mean(0) for o, k in zip(off, nats)])
    mols.set(O.M_POS, com)
    atoms = O.Atoms.from_arrays(apos, type=atype, relmass=np.array([1, 1, 0.5, 0.5, 1, 1.0]))
    vl = O.Verlet()
    vl.build(mols, 5, 0, 5, 3.5, 1.0, [-3.0] * 3, [5.0] * 3, molecules=True)
    assert vl.total() == 10
    lj = O.LJIdealGas([0.0] * 4, [rc] * 4, sig.ravel(), eps.ravel(), 2, shifted=True)
    lj.set_intervals(1, 1000)
    lj.run(mols, vl, atoms)  # run 0: samples, then the update moves them into the mean
    O.lib().or_atoms_set_force(atoms.h, 0.0)
    O.lib().or_mol_set_force(mols.h, 0.0)
    mean0 = lj.hist(2).copy()
    energy = lj.run(mols, vl, atoms)  # run 1: samples stay in the accumulators
    hy = [False, True, True, False, False]
    binof = [int(math.floor(l * 200.0)) for l in lam]
    comp = np.zeros((200, 2))
    cnt = np.zeros((200, 2))
    e_tot = 0.0
    Fa = np.zeros((6, 3))
    Fm = np.zeros((5, 3))
    for a in range(5):
        for b in range(a + 1, 5):
            if mlam[a] <= 0.0 and mlam[b] <= 0.0:
                continue
            w = 0.5 * (mlam[a] + mlam[b])
            for i in range(off[a], off[a] + nats[a]):
                for j in range(off[b], off[b] + nats[b]):
                    d = apos[i] - apos[j]
                    r2 = float(d @ d)
                    if r2 > rc * rc:
                        continue
                    ff, e = _lj_pair(r2, eps[atype[i], atype[j]], sig[atype[i], atype[j]], rc)
                    Fa[i] += w * ff * d
                    Fa[j] -= w * ff * d
                    e_tot += w * e
                    if hy[a] or hy[b]:
                        Fm[a] -= 0.5 * e * np.array(grad[a])
                        Fm[b] -= 0.5 * e * np.array(grad[b])
                        if hy[a]:
                            comp[binof[a], atype[i]] += 0.5 * e
                        if hy[b]:
                            comp[binof[b], atype[j]] += 0.5 * e
    for a in range(5):
        if hy[a]:
            for i in range(off[a], off[a] + nats[a]):
                cnt[binof[a], atype[i]] += 1.0
    # the mean after run 0 is comp / cnt / 11 where counted; run 1 adds mean * grad
    expect_mean = np.where(cnt > 0, comp / np.where(cnt > 0, cnt, 1.0) / 11.0, 0.0)
    np.testing.assert_allclose(mean0, expect_mean, rtol=1e-13, atol=1e-16)
    for a in range(5):
        if hy[a]:
            Fm[a] += expect_mean[binof[a], atype[off[a]]] * np.array(grad[a])
    assert energy == pytest.approx(e_tot, rel=1e-13)
    np.testing.assert_allclose(lj.hist(0), comp, rtol=1e-13, atol=1e-16)
    np.testing.assert_array_equal(lj.hist(1), cnt)
    np.testing.assert_allclose(atoms.get(O.FORCE, 6), Fa, rtol=1e-12, atol=1e-13)
    np.testing.assert_allclose(mols.get(O.M_FORCE, 5), Fm, rtol=1e-12, atol=1e-14)
    # the diatomic hybrid molecule counts two atoms, in its atoms' type columns
    assert cnt[binof[2]].tolist() == [1.0, 1.0]
    assert comp[binof[1], 1] != 0.0 and comp[binof[2], 0] != 0.0
