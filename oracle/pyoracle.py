"""ctypes view of the CPU oracle (TEST INFRASTRUCTURE ONLY).

Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may import this
module.  It wraps oracle/_build/liboracle.so, the C++ restatement of the reference's hot
path (see oracle/oracle.cpp for the per-function reference citations).
"""
import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "_build", "liboracle.so")


def build():
    subprocess.check_call(["make", "-s", "-C", HERE])


class Box(C.Structure):
    _fields_ = [("min", C.c_double * 3), ("max", C.c_double * 3), ("ghost", C.c_double * 3)]


class Pred1(C.Structure):
    _fields_ = [("kind", C.c_int32), ("axis", C.c_int32), ("a", C.c_double), ("b", C.c_double),
                ("c", C.c_double), ("tol", C.c_double)]


class Pred2(C.Structure):
    _fields_ = [("op", C.c_int32), ("pad", C.c_int32), ("first", Pred1), ("second", Pred1)]


class Slab(C.Structure):
    _fields_ = [("center", C.c_double * 3), ("atomistic_diameter", C.c_double),
                ("hybrid_diameter", C.c_double), ("nu", C.c_int64), ("interface_type", C.c_int32),
                ("pad", C.c_int32)]


class MDParams(C.Structure):
    _fields_ = [("dt", C.c_double), ("skin", C.c_double), ("rc", C.c_double), ("sigma", C.c_double),
                ("epsilon", C.c_double), ("cap", C.c_double), ("cell_ratio", C.c_double),
                ("gamma", C.c_double), ("temperature", C.c_double), ("seed", C.c_uint64),
                ("max_neigh", C.c_int64), ("nsteps", C.c_int64), ("energy_every", C.c_int64),
                ("shifted", C.c_int32), ("langevin", C.c_int32), ("permute", C.c_int32),
                ("pad", C.c_int32)]


def make_box(lo, hi, ghost):
    b = Box()
    for d in range(3):
        b.min[d] = lo[d]
        b.max[d] = hi[d]
        b.ghost[d] = ghost if np.isscalar(ghost) else ghost[d]
    return b


def make_slab(center, at_diam, hy_diam, nu, abrupt=False):
    s = Slab()
    for d in range(3):
        s.center[d] = center[d]
    s.atomistic_diameter = at_diam
    s.hybrid_diameter = hy_diam
    s.nu = nu
    s.interface_type = 1 if abrupt else 0
    return s


_lib = None
P = C.c_void_p
D = C.c_double
I64 = C.c_int64
I32 = C.c_int
DP = C.POINTER(C.c_double)


def lib():
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        build()
    L = C.CDLL(LIB_PATH)
    sig = {
        "or_set_threads": (None, [I32]),
        "or_get_threads": (I32, []),
        "or_philox4x32_10": (None, [P, P, P]),
        "or_det_log": (D, [D]),
        "or_det_exp": (D, [D]),
        "or_normals3": (None, [C.c_uint64, C.c_uint64, C.c_uint64, P]),
        "or_atoms_new": (P, []),
        "or_atoms_free": (None, [P]),
        "or_atoms_resize": (None, [P, I64]),
        "or_atoms_size": (I64, [P]),
        "or_atoms_set_counts": (None, [P, I64, I64]),
        "or_atoms_nlocal": (I64, [P]),
        "or_atoms_nghost": (I64, [P]),
        "or_atoms_set": (None, [P, I32, P, I64]),
        "or_atoms_get": (None, [P, I32, P, I64]),
        "or_atoms_set_force": (None, [P, D]),
        "or_mol_new": (P, []),
        "or_mol_free": (None, [P]),
        "or_mol_resize": (None, [P, I64]),
        "or_mol_size": (I64, [P]),
        "or_mol_set_counts": (None, [P, I64, I64]),
        "or_mol_nlocal": (I64, [P]),
        "or_mol_nghost": (I64, [P]),
        "or_mol_set": (None, [P, I32, P, I64]),
        "or_mol_get": (None, [P, I32, P, I64]),
        "or_mol_set_force": (None, [P, D]),
        "or_create_molecule_for_each_atom": (P, [P]),
        "or_ghost_new": (P, []),
        "or_ghost_free": (None, [P]),
        "or_ghost_get_corr": (None, [P, P, I64]),
        "or_ghost_corr_size": (I64, [P]),
        "or_ghost_set_corr": (None, [P, P, I64]),
        "or_map_into_domain": (None, [P, P]),
        "or_ghost_reset": (None, [P, P]),
        "or_ghost_create_axis": (None, [P, P, P, I32]),
        "or_ghost_create_xyz": (None, [P, P, P]),
        "or_ghost_update": (None, [P, P, P]),
        "or_ghost_fold": (None, [P, P]),
        "or_cells_build": (P, [P, I64, I64, P, P, P]),
        "or_cells_free": (None, [P]),
        "or_cells_dims": (None, [P, P]),
        "or_cells_get": (None, [P, P, P, P, P]),
        "or_atoms_permute": (None, [P, P]),
        "or_verlet_new": (P, []),
        "or_verlet_free": (None, [P]),
        "or_verlet_build": (None, [P, P, I64, I64, I64, D, D, P, P, I64, I32]),
        "or_verlet_build_mol": (None, [P, P, I64, I64, I64, D, D, P, P, I64, I32]),
        "or_verlet_maxn": (I64, [P]),
        "or_verlet_nrows": (I64, [P]),
        "or_verlet_counts": (None, [P, P]),
        "or_verlet_neighbors": (None, [P, P]),
        "or_verlet_total": (I64, [P]),
        "or_verlet_grid": (None, [P, P, P]),
        "or_lj_new": (P, [P, P, P, P, I64, I32, I64]),
        "or_lj_free": (None, [P]),
        "or_lj_table": (None, [P, P]),
        "or_lj_eval": (None, [P, D, I64, P, P]),
        "or_lj_apply": (None, [P, P, P, P]),
        "or_lj_energy": (D, [P]),
        "or_lj_virial": (D, [P]),
        "or_vv_pre": (D, [P, D]),
        "or_vv_post": (None, [P, D]),
        "or_langevin_pre": (D, [P, D, D, D, C.c_uint64, C.c_uint64, P]),
        "or_kinetic_energy": (D, [P]),
        "or_slab_eval": (None, [P, D, P]),
        "or_spherical_eval": (None, [P, D, D, I64, D, D, D, P]),
        "or_berendsen_apply": (I32, [P, D, D, D]),
        "or_update_molecules": (None, [P, P, P]),
        "or_contribute_molecule_force": (None, [P, P]),
        "or_multires_exchange_real": (None, [P, P, P]),
        "or_multires_ghost_axis": (None, [P, P, P, P, I32]),
        "or_multires_ghost_xyz": (None, [P, P, P, P]),
        "or_ljig_new": (P, [P, P, P, P, I64, I32]),
        "or_ljig_free": (None, [P]),
        "or_ljig_set_intervals": (None, [P, I64, I64]),
        "or_ljig_get_hist": (None, [P, I32, P]),
        "or_ljig_set_mean": (None, [P, P]),
        "or_ljig_run": (D, [P, P, P, P]),
        "or_tf_new": (P, [P, P, I64, P, D, I32, I32]),
        "or_tf_free": (None, [P]),
        "or_tf_num_bins": (I64, [P]),
        "or_tf_samples": (I64, [P]),
        "or_tf_get": (None, [P, I32, P]),
        "or_tf_set_force": (None, [P, P]),
        "or_tf_sample": (None, [P, P]),
        "or_tf_update": (I32, [P, D, D, P]),
        "or_tf_apply": (None, [P, P, P]),
        "or_tf_apply_interpolated": (None, [P, P, P]),
        "or_tf_mu": (None, [P, P, P]),
        "or_hist_gradient": (None, [D, D, I64, I64, P, I32, P]),
        "or_hist_smoothen": (None, [D, D, I64, I64, P, D, D, I32, P]),
        "or_hist_make_symmetric": (None, [I64, I64, P]),
        "or_hist_get_bin": (I64, [D, D, I64, D]),
        "or_hist_bin_position": (D, [D, D, I64, I64]),
        "or_run_md": (I64, [P, P, P, P]),
        "or_lj_adjoint": (None, [P, P, P, P, P, P, P, P]),
        "or_md_new": (P, [P, P, P]),
        "or_md_free": (None, [P]),
        "or_md_steps": (I64, [P, I64]),
        "or_md_energy": (D, [P]),
    }
    for name, (res, args) in sig.items():
        f = getattr(L, name)
        f.restype = res
        f.argtypes = args
    _lib = L
    return L


def ptr(a):
    return a.ctypes.data_as(C.c_void_p) if a is not None else None


def ref(s):
    return C.byref(s) if s is not None else None


# field ids
POS, VEL, FORCE, TYPE, MASS, CHARGE, RELMASS = range(7)
M_POS, M_FORCE, M_LAMBDA, M_MODLAMBDA, M_GRAD, M_OFFSET, M_NUMATOMS = range(7)
_VEC_A = {POS, VEL, FORCE}
_INT_A = {TYPE}
_VEC_M = {M_POS, M_FORCE, M_GRAD}
_INT_M = {M_OFFSET, M_NUMATOMS}


class Atoms:
    """AoS atoms record array owned by the oracle (mirror of data::Atoms)."""

    def __init__(self, n=0):
        self.h = lib().or_atoms_new()
        if n:
            lib().or_atoms_resize(self.h, n)

    def __del__(self):
        if getattr(self, "h", None):
            lib().or_atoms_free(self.h)
            self.h = None

    @property
    def size(self):
        return lib().or_atoms_size(self.h)

    @property
    def nlocal(self):
        return lib().or_atoms_nlocal(self.h)

    @property
    def nghost(self):
        return lib().or_atoms_nghost(self.h)

    def set_counts(self, nl, ng=0):
        lib().or_atoms_set_counts(self.h, nl, ng)

    def set(self, field, arr):
        if field in _VEC_A:
            a = np.ascontiguousarray(arr, dtype=np.float64).reshape(-1, 3)
        elif field in _INT_A:
            a = np.ascontiguousarray(arr, dtype=np.int64).reshape(-1)
        else:
            a = np.ascontiguousarray(arr, dtype=np.float64).reshape(-1)
        lib().or_atoms_set(self.h, field, ptr(a), a.shape[0])

    def get(self, field, n=None):
        n = self.size if n is None else n
        if field in _VEC_A:
            out = np.zeros((n, 3))
        elif field in _INT_A:
            out = np.zeros(n, dtype=np.int64)
        else:
            out = np.zeros(n)
        lib().or_atoms_get(self.h, field, ptr(out), n)
        return out

    @classmethod
    def from_arrays(cls, pos, vel=None, force=None, type=None, mass=None, relmass=None, nghost=0):
        pos = np.asarray(pos, dtype=np.float64).reshape(-1, 3)
        n = pos.shape[0]
        a = cls(n)
        a.set(POS, pos)
        a.set(VEL, np.zeros((n, 3)) if vel is None else vel)
        a.set(FORCE, np.zeros((n, 3)) if force is None else force)
        a.set(TYPE, np.zeros(n, np.int64) if type is None else type)
        a.set(MASS, np.ones(n) if mass is None else mass)
        a.set(RELMASS, np.ones(n) if relmass is None else relmass)
        a.set_counts(n - nghost, nghost)
        return a


class Molecules:
    def __init__(self, n=0, handle=None):
        self.h = handle if handle is not None else lib().or_mol_new()
        if n:
            lib().or_mol_resize(self.h, n)

    def __del__(self):
        if getattr(self, "h", None):
            lib().or_mol_free(self.h)
            self.h = None

    @property
    def size(self):
        return lib().or_mol_size(self.h)

    @property
    def nlocal(self):
        return lib().or_mol_nlocal(self.h)

    @property
    def nghost(self):
        return lib().or_mol_nghost(self.h)

    def set_counts(self, nl, ng=0):
        lib().or_mol_set_counts(self.h, nl, ng)

    def set(self, field, arr):
        if field in _VEC_M:
            a = np.ascontiguousarray(arr, dtype=np.float64).reshape(-1, 3)
        elif field in _INT_M:
            a = np.ascontiguousarray(arr, dtype=np.int64).reshape(-1)
        else:
            a = np.ascontiguousarray(arr, dtype=np.float64).reshape(-1)
        lib().or_mol_set(self.h, field, ptr(a), a.shape[0])

    def get(self, field, n=None):
        n = self.size if n is None else n
        if field in _VEC_M:
            out = np.zeros((n, 3))
        elif field in _INT_M:
            out = np.zeros(n, dtype=np.int64)
        else:
            out = np.zeros(n)
        lib().or_mol_get(self.h, field, ptr(out), n)
        return out

    @classmethod
    def for_each_atom(cls, atoms):
        return cls(handle=lib().or_create_molecule_for_each_atom(atoms.h))


class Ghost:
    def __init__(self):
        self.h = lib().or_ghost_new()

    def __del__(self):
        if getattr(self, "h", None):
            lib().or_ghost_free(self.h)
            self.h = None

    def corr(self, n=None):
        n = lib().or_ghost_corr_size(self.h) if n is None else n
        out = np.zeros(n, dtype=np.int64)
        lib().or_ghost_get_corr(self.h, ptr(out), n)
        return out

    def set_corr(self, c):
        c = np.ascontiguousarray(c, dtype=np.int64)
        lib().or_ghost_set_corr(self.h, ptr(c), c.size)

    def create_xyz(self, atoms, box):
        lib().or_ghost_create_xyz(self.h, atoms.h, C.byref(box))

    def create_axis(self, atoms, box, axis):
        lib().or_ghost_create_axis(self.h, atoms.h, C.byref(box), axis)

    def reset(self, atoms):
        lib().or_ghost_reset(self.h, atoms.h)

    def update(self, atoms, box):
        lib().or_ghost_update(self.h, atoms.h, C.byref(box))

    def fold(self, atoms):
        lib().or_ghost_fold(self.h, atoms.h)


class Verlet:
    def __init__(self):
        self.h = lib().or_verlet_new()

    def __del__(self):
        if getattr(self, "h", None):
            lib().or_verlet_free(self.h)
            self.h = None

    def build(self, atoms, nall, begin, end, rn, ratio, gmin, gmax, max_neigh=0, half=True,
              molecules=False):
        gmin = np.ascontiguousarray(gmin, dtype=np.float64)
        gmax = np.ascontiguousarray(gmax, dtype=np.float64)
        f = lib().or_verlet_build_mol if molecules else lib().or_verlet_build
        f(self.h, atoms.h, nall, begin, end, rn, ratio, ptr(gmin), ptr(gmax), max_neigh,
          1 if half else 0)

    def counts(self):
        n = lib().or_verlet_nrows(self.h)
        out = np.zeros(n, dtype=np.int32)
        lib().or_verlet_counts(self.h, ptr(out))
        return out

    def neighbors(self):
        n = lib().or_verlet_nrows(self.h)
        m = lib().or_verlet_maxn(self.h)
        out = np.zeros((n, max(m, 0)), dtype=np.int32)
        if n * m:
            lib().or_verlet_neighbors(self.h, ptr(out))
        return out

    def rows(self, begin, end):
        """list of neighbour sets per row"""
        cnt = self.counts()
        nb = self.neighbors()
        return [set(nb[i, :cnt[i]].tolist()) for i in range(begin, end)]

    def total(self):
        return lib().or_verlet_total(self.h)

    def grid(self):
        n = np.zeros(3, dtype=np.int32)
        dx = np.zeros(3)
        lib().or_verlet_grid(self.h, ptr(n), ptr(dx))
        return n, dx


class Cells:
    def __init__(self, atoms, begin, end, delta, gmin, gmax):
        delta = np.ascontiguousarray(delta, dtype=np.float64)
        gmin = np.ascontiguousarray(gmin, dtype=np.float64)
        gmax = np.ascontiguousarray(gmax, dtype=np.float64)
        self.begin, self.end = begin, end
        self.h = lib().or_cells_build(atoms.h, begin, end, ptr(delta), ptr(gmin), ptr(gmax))

    def __del__(self):
        if getattr(self, "h", None):
            lib().or_cells_free(self.h)
            self.h = None

    def dims(self):
        n = np.zeros(3, dtype=np.int32)
        lib().or_cells_dims(self.h, ptr(n))
        return n

    def get(self):
        n = self.dims()
        nc = int(np.prod(n))
        m = self.end - self.begin
        cell = np.zeros(m, dtype=np.int32)
        counts = np.zeros(nc, dtype=np.int32)
        offsets = np.zeros(nc, dtype=np.int64)
        perm = np.zeros(m, dtype=np.int64)
        lib().or_cells_get(self.h, ptr(cell), ptr(counts), ptr(offsets), ptr(perm))
        return cell, counts, offsets, perm

    def permute(self, atoms):
        lib().or_atoms_permute(atoms.h, self.h)


class LJ:
    def __init__(self, cap, rc, sigma, eps, ntypes=1, shifted=False, type_stride=None):
        arr = lambda v: np.ascontiguousarray(np.atleast_1d(v), dtype=np.float64)
        self._keep = [arr(cap), arr(rc), arr(sigma), arr(eps)]
        ts = 1 if type_stride is None else type_stride
        self.h = lib().or_lj_new(*[ptr(a) for a in self._keep], ntypes, 1 if shifted else 0, ts)
        self.ntypes = ntypes

    def __del__(self):
        if getattr(self, "h", None):
            lib().or_lj_free(self.h)
            self.h = None

    def table(self):
        out = np.zeros((self.ntypes * self.ntypes, 10))
        lib().or_lj_table(self.h, ptr(out))
        return out

    def eval(self, dist_sqr, type_idx=0):
        ff = C.c_double()
        e = C.c_double()
        lib().or_lj_eval(self.h, dist_sqr, type_idx, C.byref(ff), C.byref(e))
        return ff.value, e.value

    def apply(self, atoms, verlet, pred=None):
        lib().or_lj_apply(self.h, atoms.h, verlet.h, ref(pred))
        return lib().or_lj_energy(self.h), lib().or_lj_virial(self.h)


class LJIdealGas:
    def __init__(self, cap, rc, sigma, eps, ntypes=1, shifted=True):
        arr = lambda v: np.ascontiguousarray(np.atleast_1d(v), dtype=np.float64)
        self._keep = [arr(cap), arr(rc), arr(sigma), arr(eps)]
        self.h = lib().or_ljig_new(*[ptr(a) for a in self._keep], ntypes, 1 if shifted else 0)
        self.ntypes = ntypes

    def __del__(self):
        if getattr(self, "h", None):
            lib().or_ljig_free(self.h)
            self.h = None

    def set_intervals(self, sampling, update):
        lib().or_ljig_set_intervals(self.h, sampling, update)

    def hist(self, which):
        out = np.zeros((200, self.ntypes))
        lib().or_ljig_get_hist(self.h, which, ptr(out))
        return out

    def set_mean(self, m):
        m = np.ascontiguousarray(m, dtype=np.float64)
        lib().or_ljig_set_mean(self.h, ptr(m))

    def run(self, molecules, verlet, atoms):
        return lib().or_ljig_run(self.h, molecules.h, verlet.h, atoms.h)


class ThermodynamicForce:
    def __init__(self, target, box, bin_width, modulation, sym=False, periodic=False):
        t = np.ascontiguousarray(np.atleast_1d(target), dtype=np.float64)
        m = np.ascontiguousarray(np.atleast_1d(modulation), dtype=np.float64)
        self.ntypes = t.size
        self.h = lib().or_tf_new(ptr(t), ptr(m), t.size, C.byref(box), bin_width, int(sym),
                                 int(periodic))

    def __del__(self):
        if getattr(self, "h", None):
            lib().or_tf_free(self.h)
            self.h = None

    @property
    def nbins(self):
        return lib().or_tf_num_bins(self.h)

    @property
    def samples(self):
        return lib().or_tf_samples(self.h)

    def force(self):
        out = np.zeros((self.nbins, self.ntypes))
        lib().or_tf_get(self.h, 0, ptr(out))
        return out

    def density(self):
        out = np.zeros((self.nbins, self.ntypes))
        lib().or_tf_get(self.h, 1, ptr(out))
        return out

    def set_force(self, f):
        f = np.ascontiguousarray(f, dtype=np.float64)
        lib().or_tf_set_force(self.h, ptr(f))

    def sample(self, atoms):
        lib().or_tf_sample(self.h, atoms.h)

    def update(self, sigma, rng, pred=None):
        return lib().or_tf_update(self.h, sigma, rng, ref(pred))

    def apply(self, atoms, pred=None):
        lib().or_tf_apply(self.h, atoms.h, ref(pred))

    def apply_interpolated(self, atoms, pred=None):
        lib().or_tf_apply_interpolated(self.h, atoms.h, ref(pred))

    def mu(self):
        l = np.zeros(self.ntypes)
        r = np.zeros(self.ntypes)
        lib().or_tf_mu(self.h, ptr(l), ptr(r))
        return l, r


def pred1(kind=0, axis=0, a=0.0, b=0.0, c=0.0, tol=0.0):
    p = Pred1()
    p.kind, p.axis, p.a, p.b, p.c, p.tol = kind, axis, a, b, c, tol
    return p


def pred2(op, first, second=None):
    p = Pred2()
    p.op = op
    p.first = first
    p.second = first if second is None else second
    return p


def md_params(**kw):
    p = MDParams()
    defaults = dict(dt=0.005, skin=0.3, rc=2.5, sigma=1.0, epsilon=1.0, cap=0.0, cell_ratio=1.0,
                    gamma=1.0, temperature=1.0, seed=1234, max_neigh=0, nsteps=1, energy_every=0,
                    shifted=1, langevin=0, permute=0)
    defaults.update(kw)
    for k, v in defaults.items():
        setattr(p, k, v)
    return p


PRED_SLAB_HY = 5


class AdResSRun:
    """the AdResS step of a monatomic LJ / ideal-gas slab through the oracle, in the order of
    the single-system driver (Langevin pre-step with the skin criterion, multi-resolution
    exchange + ghosts + molecule list on rebuild steps, ghost update otherwise, molecule
    update with the slab, LJ_IdealGas, molecule forces to atoms, ghost fold, density sample,
    thermodynamic-force update every tf_every steps over the hybrid region, its application,
    post-step) -- the loop of examples/04_LennardJones_IdealGas_LocalCap"""

    def __init__(self, pos, vel, L, atomistic_diameter, hybrid_diameter, nu=1, rc=2.5, skin=0.3, cap=0.0,
                 gamma=1.0, temperature=1.0, seed=1234, tf_every=100, tf_sigma=2.0, tf_range=3.0, shifted=True):
        self.L = [float(x) for x in L]
        self.rn, self.skin = rc + skin, skin
        self.gamma, self.temperature, self.seed = gamma, temperature, seed
        self.tf_every, self.tf_sigma, self.tf_range = tf_every, tf_sigma, tf_range
        self.atoms = Atoms.from_arrays(np.asarray(pos, dtype=np.float64), vel=np.asarray(vel, dtype=np.float64))
        self.mols = Molecules.for_each_atom(self.atoms)
        self.box = make_box([0, 0, 0], self.L, self.rn)
        c = [0.5 * x for x in self.L]
        self.slab = make_slab(c, atomistic_diameter, hybrid_diameter, nu)
        self.ghost, self.verlet = Ghost(), Verlet()
        self.lj = LJIdealGas([cap], [rc], [1.0], [1.0], 1, shifted=shifted)
        rho = pos.shape[0] / float(np.prod(self.L))
        self.tf = ThermodynamicForce([rho], self.box, 1.0, [1.0], sym=True, periodic=True)
        self.hy = pred1(kind=PRED_SLAB_HY, axis=0, a=c[0], b=0.5 * atomistic_diameter, c=hybrid_diameter)
        self.lo, self.hi = [-self.rn] * 3, [x + self.rn for x in self.L]
        self.disp, self.rebuilds, self.calls, self.steps = np.inf, 0, 0, 0

    def _rebuild(self):
        a, m = self.atoms, self.mols
        lib().or_update_molecules(m.h, a.h, None)
        lib().or_multires_exchange_real(C.byref(self.box), m.h, a.h)
        lib().or_multires_ghost_xyz(self.ghost.h, m.h, a.h, C.byref(self.box))
        lib().or_update_molecules(m.h, a.h, None)
        self.verlet.build(m, m.nlocal + m.nghost, 0, m.nlocal, self.rn, 1.0, self.lo, self.hi, molecules=True)
        self.rebuilds += 1

    def step(self, dt):
        a, m = self.atoms, self.mols
        self.disp += lib().or_langevin_pre(a.h, dt, self.gamma, self.temperature, self.seed, self.calls, None)
        self.calls += 1
        if self.disp >= 0.5 * self.skin:
            self.disp = 0.0
            self._rebuild()
        else:
            self.ghost.update(a, self.box)
        lib().or_update_molecules(m.h, a.h, C.byref(self.slab))
        lib().or_atoms_set_force(a.h, 0.0)
        lib().or_mol_set_force(m.h, 0.0)
        energy = self.lj.run(m, self.verlet, a)
        lib().or_contribute_molecule_force(m.h, a.h)
        self.ghost.fold(a)
        self.tf.sample(a)
        self.steps += 1
        if self.tf_every and self.steps % self.tf_every == 0:
            self.tf.update(self.tf_sigma, self.tf_range, self.hy)
        self.tf.apply(a)
        lib().or_vv_post(a.h, dt)
        return energy


def read_gro(filename):
    """io::restoreGRO's reading of a GROMACS coordinate file (io/RestoreGRO.cpp:24-147), here
    by the format's fixed columns: atom lines hold the position in columns 21-44 (three 8-wide
    fields) and the velocity in columns 45-68; the line after the atoms holds the box edges.
    Returns (positions, velocities, edges)."""
    with open(filename) as f:
        lines = f.read().splitlines()
    n = int(lines[1].strip())
    pos = np.zeros((n, 3))
    vel = np.zeros((n, 3))
    for i in range(n):
        ln = lines[2 + i]
        pos[i] = [float(ln[20 + 8 * d:28 + 8 * d]) for d in range(3)]
        if len(ln.rstrip()) > 44:
            vel[i] = [float(ln[44 + 8 * d:52 + 8 * d]) for d in range(3)]
    edges = np.array([float(w) for w in lines[2 + n].split()[:3]])
    return pos, vel, edges
