"""Python view of the gfx950 MRMD hot-path library (mirrors the reference's C++ action API).

The classes carry the reference's names (data::Atoms, communication::GhostLayer,
HalfVerletList, action::LennardJones, ...) and call the C ABI of include/mrmd_hip.h
through ctypes.  Device memory belongs to the library; host data moves through
upload/download.  Importing fails loudly when libmrmd_hip.so is missing: there is no CPU
fallback.
"""
import ctypes as C

import numpy as np

from . import capi
from .capi import Box, MrmdError, Pred1, Pred2, Slab as SlabParams, Spherical as SphericalParams

_L = capi.lib

POS, VEL, FORCE, TYPE, MASS, CHARGE, RELMASS = range(7)
M_POS, M_FORCE, M_LAMBDA, M_MODLAMBDA, M_GRAD, M_OFFSET, M_NUMATOMS = range(7)
_VEC = {POS, VEL, FORCE}
_M_VEC = {M_POS, M_FORCE, M_GRAD}
_M_INT = {M_OFFSET, M_NUMATOMS}

PRED_ALWAYS, PRED_NEVER, PRED_OPEN_RANGE, PRED_SYMMETRIC_SLAB = 0, 1, 2, 3
PRED_SLAB_AT, PRED_SLAB_HY, PRED_SLAB_CG, PRED_GREATER, PRED_LESS = 4, 5, 6, 7, 8
PAIR_ALWAYS, PAIR_AND, PAIR_OR = 0, 1, 2


def _vec3(v):
    """a per-axis triple of doubles (a scalar applies to all three axes); the library reads
    exactly three"""
    a = np.asarray(v, dtype=np.float64)
    if a.size == 1:
        a = np.full(3, float(a.reshape(-1)[0]))
    if a.shape != (3,):
        raise ValueError(f"expected 3 values per axis, got shape {a.shape}")
    return np.ascontiguousarray(a)


def _p(a):
    return a.ctypes.data_as(C.c_void_p)


def _i64(v):
    return C.c_int64(int(v))


class Context:
    """one device + one HIP stream (mrmd_ctx)"""

    def __init__(self, device=0):
        h = C.c_void_p()
        rc = _L.mrmd_ctx_create(device, C.byref(h))
        if rc != 0:
            raise MrmdError(f"mrmd_ctx_create failed ({rc}); is this a gfx950 device?")
        self.h = h

    def check(self, rc):
        if rc != 0:
            raise MrmdError(f"[{rc}] {_L.mrmd_ctx_last_error(self.h).decode()}")

    def close(self):
        if getattr(self, "h", None):
            _L.mrmd_ctx_destroy(self.h)
            self.h = None

    def __del__(self):
        self.close()

    def synchronize(self):
        self.check(_L.mrmd_ctx_synchronize(self.h))

    def set_stream(self, stream_handle):
        self.check(_L.mrmd_ctx_set_stream(self.h, C.c_void_p(stream_handle)))

    profiling = False

    def set_profiling(self, on=True, only=None):
        """per-kernel HIP-event timing; only = one timer name (the others record nothing)"""
        self.check(_L.mrmd_ctx_set_profile_filter(self.h, only.encode() if only else None))
        self.check(_L.mrmd_ctx_set_profiling(self.h, 1 if on else 0))
        self.profiling = bool(on)

    def kernel_stats(self, name):
        n = C.c_int64()
        ms = C.c_double()
        self.check(_L.mrmd_ctx_kernel_stats(self.h, name.encode(), C.byref(n), C.byref(ms)))
        return n.value, ms.value

    def reset_stats(self):
        self.check(_L.mrmd_ctx_reset_stats(self.h))

    def device_bytes(self):
        b = C.c_int64()
        self.check(_L.mrmd_ctx_device_bytes(self.h, C.byref(b)))
        return b.value


def make_box(lo, hi, ghost):
    b = Box()
    g = [ghost] * 3 if np.isscalar(ghost) else list(ghost)
    for d in range(3):
        b.min[d], b.max[d], b.ghost[d] = lo[d], hi[d], g[d]
    return b


def pred1(kind=PRED_ALWAYS, axis=0, a=0.0, b=0.0, c=0.0, tol=0.0):
    p = Pred1()
    p.kind, p.axis, p.a, p.b, p.c, p.tol = kind, axis, a, b, c, tol
    return p


def pred2(op, first, second=None):
    p = Pred2()
    p.op = op
    p.first = first
    p.second = first if second is None else second
    return p


class Slab:
    """weighting_function::Slab (Slab.hpp:189-200)"""

    def __init__(self, center, atomistic_diameter, hybrid_diameter, nu, abrupt=False):
        self.params = SlabParams()
        for d in range(3):
            self.params.center[d] = center[d]
        self.params.atomistic_diameter = atomistic_diameter
        self.params.hybrid_diameter = hybrid_diameter
        self.params.nu = nu
        self.params.interface_type = 1 if abrupt else 0

    def __call__(self, ctx, x):
        x = np.ascontiguousarray(np.atleast_1d(x), dtype=np.float64)
        out = np.zeros((x.size, 5))
        ctx.check(_L.mrmd_slab_eval(ctx.h, C.byref(self.params), _p(x), x.size, _p(out)))
        return out


class Spherical:
    """weighting_function::Spherical(center, atomisticRadius, hybridRegionDiameter, exponent)
    (Spherical.hpp:25-99): lambda = cos^exponent over the hybrid shell around the centre"""

    def __init__(self, center, atomistic_radius, hybrid_diameter, exponent):
        self.params = SphericalParams()
        for d in range(3):
            self.params.center[d] = center[d]
        self.params.atomistic_radius = atomistic_radius
        self.params.hybrid_diameter = hybrid_diameter
        self.params.exponent = int(exponent)

    def __call__(self, ctx, xyz):
        """(n, 4): lambda and its gradient at n points"""
        x = np.ascontiguousarray(np.atleast_2d(xyz), dtype=np.float64)
        out = np.zeros((x.shape[0], 4))
        ctx.check(_L.mrmd_spherical_eval(ctx.h, C.byref(self.params), _p(x), x.shape[0], _p(out)))
        return out


class BerendsenThermostat:
    """action::BerendsenThermostat (BerendsenThermostat.hpp:22-36)"""

    @staticmethod
    def apply(atoms, current_temperature, target_temperature, gamma):
        atoms.ctx.check(_L.mrmd_berendsen_apply(atoms.h, float(current_temperature), float(target_temperature),
                                                float(gamma)))


class _Container:
    _prefix = ""
    _vec = set()
    _int = set()

    def __init__(self, ctx, n=0):
        self.ctx = ctx
        h = C.c_void_p()
        ctx.check(getattr(_L, f"mrmd_{self._prefix}_create")(ctx.h, n, C.byref(h)))
        self.h = h

    def __del__(self):
        if getattr(self, "h", None):
            getattr(_L, f"mrmd_{self._prefix}_destroy")(self.h)
            self.h = None

    def counts(self):
        nl, ng, sz = C.c_int64(), C.c_int64(), C.c_int64()
        self.ctx.check(getattr(_L, f"mrmd_{self._prefix}_get_counts")(self.h, C.byref(nl), C.byref(ng),
                                                                         C.byref(sz)))
        return nl.value, ng.value, sz.value

    @property
    def nlocal(self):
        return self.counts()[0]

    @property
    def nghost(self):
        return self.counts()[1]

    @property
    def size(self):
        return self.counts()[2]

    def set_counts(self, nl, ng=0):
        self.ctx.check(getattr(_L, f"mrmd_{self._prefix}_set_counts")(self.h, nl, ng))

    def resize(self, n):
        self.ctx.check(getattr(_L, f"mrmd_{self._prefix}_resize")(self.h, n))

    def upload(self, field, arr, offset=0):
        if field in self._vec:
            a = np.ascontiguousarray(arr, dtype=np.float64).reshape(-1, 3)
        elif field in self._int:
            a = np.ascontiguousarray(arr, dtype=np.int64).reshape(-1)
        else:
            a = np.ascontiguousarray(arr, dtype=np.float64).reshape(-1)
        self.ctx.check(getattr(_L, f"mrmd_{self._prefix}_upload")(self.h, field, _p(a), offset, a.shape[0]))

    def download(self, field, n=None, offset=0):
        n = self.size - offset if n is None else n
        if field in self._vec:
            out = np.zeros((n, 3))
        elif field in self._int:
            out = np.zeros(n, dtype=np.int64)
        else:
            out = np.zeros(n)
        if n:
            self.ctx.check(getattr(_L, f"mrmd_{self._prefix}_download")(self.h, field, _p(out), offset, n))
        return out

    def fill(self, field, value):
        self.ctx.check(getattr(_L, f"mrmd_{self._prefix}_fill")(self.h, field, float(value)))


class Atoms(_Container):
    """data::Atoms (Atoms.hpp:33-177) in device memory"""
    _prefix = "atoms"
    _vec = _VEC
    _int = {TYPE}

    @classmethod
    def from_arrays(cls, ctx, pos, vel=None, force=None, type=None, mass=None, relmass=None, charge=None,
                    nghost=0):
        pos = np.asarray(pos, dtype=np.float64).reshape(-1, 3)
        n = pos.shape[0]
        a = cls(ctx, n)
        a.upload(POS, pos)
        if vel is not None:
            a.upload(VEL, vel)
        if force is not None:
            a.upload(FORCE, force)
        a.upload(TYPE, np.zeros(n, np.int64) if type is None else type)
        a.upload(MASS, np.ones(n) if mass is None else mass)
        a.upload(RELMASS, np.ones(n) if relmass is None else relmass)
        if charge is not None:
            a.upload(CHARGE, charge)
        a.set_counts(n - nghost, nghost)
        return a

    def set_force(self, value):
        self.fill(FORCE, value)

    def permute(self, cell_list):
        self.ctx.check(_L.mrmd_atoms_permute(self.h, cell_list.h))

    def kinetic_energy(self):
        e = C.c_double()
        self.ctx.check(_L.mrmd_kinetic_energy(self.h, C.byref(e)))
        return e.value


class Molecules(_Container):
    """data::Molecules (Molecules.hpp:27-176) in device memory"""
    _prefix = "molecules"
    _vec = _M_VEC
    _int = _M_INT

    @classmethod
    def for_each_atom(cls, ctx, atoms):
        m = cls(ctx, 0)
        ctx.check(_L.mrmd_molecules_for_each_atom(m.h, atoms.h))
        return m

    def update(self, atoms, slab=None):
        self.ctx.check(_L.mrmd_molecules_update(self.h, atoms.h, C.byref(slab.params) if slab else None))

    def contribute_force(self, atoms):
        self.ctx.check(_L.mrmd_molecules_contribute_force(self.h, atoms.h))

    def set_force(self, value):
        self.fill(M_FORCE, value)


class GhostLayer:
    """communication::GhostLayer (GhostLayer.hpp:28-56)"""

    def __init__(self, ctx):
        self.ctx = ctx
        h = C.c_void_p()
        ctx.check(_L.mrmd_ghost_layer_create(ctx.h, C.byref(h)))
        self.h = h

    def __del__(self):
        if getattr(self, "h", None):
            _L.mrmd_ghost_layer_destroy(self.h)
            self.h = None

    def exchange_real_atoms(self, atoms, box):
        self.ctx.check(_L.mrmd_halo_wrap(atoms.h, C.byref(box)))

    def create_ghost_atoms(self, atoms, box):
        n = C.c_int64()
        self.ctx.check(_L.mrmd_halo_create(self.h, atoms.h, C.byref(box), C.byref(n)))
        return n.value

    def create_ghost_atoms_axis(self, atoms, box, axis, reset=False):
        self.ctx.check(_L.mrmd_halo_create_axis(self.h, atoms.h, C.byref(box), axis, 1 if reset else 0))

    def update_ghost_atoms(self, atoms, box):
        self.ctx.check(_L.mrmd_halo_update(self.h, atoms.h, C.byref(box)))

    def contribute_back_ghost_to_real(self, atoms):
        self.ctx.check(_L.mrmd_halo_fold_forces(self.h, atoms.h))

    def corresponding(self, n):
        out = np.zeros(n, dtype=np.int64)
        if n:
            self.ctx.check(_L.mrmd_halo_get_corresponding(self.h, _p(out), n))
        return out

    def set_corresponding(self, c):
        c = np.ascontiguousarray(c, dtype=np.int64)
        self.ctx.check(_L.mrmd_halo_set_corresponding(self.h, _p(c), c.size))

    # communication::MultiResGhostLayer (MultiResGhostLayer.hpp:29-61)
    def multires_exchange_real_atoms(self, molecules, atoms, box):
        self.ctx.check(_L.mrmd_multires_wrap(molecules.h, atoms.h, C.byref(box)))

    def multires_create_ghost_atoms(self, molecules, atoms, box):
        self.ctx.check(_L.mrmd_multires_halo_create(self.h, molecules.h, atoms.h, C.byref(box)))

    def multires_create_ghost_atoms_axis(self, molecules, atoms, box, axis, reset=False):
        self.ctx.check(_L.mrmd_multires_halo_create_axis(self.h, molecules.h, atoms.h, C.byref(box), axis,
                                                         1 if reset else 0))


class LinkedCellList:
    """Cabana::LinkedCellList(pos, begin, end, delta, min, max) with stable in-cell order"""

    def __init__(self, ctx, atoms, begin, end, delta, gmin, gmax):
        self.ctx = ctx
        h = C.c_void_p()
        ctx.check(_L.mrmd_cell_list_create(ctx.h, C.byref(h)))
        self.h = h
        self.begin, self.end = begin, end
        self.build(atoms, begin, end, delta, gmin, gmax)

    def build(self, atoms, begin, end, delta, gmin, gmax):
        """rebuild in place (buffers are reused)"""
        self.begin, self.end = begin, end
        d = _vec3(delta)
        lo, hi = _vec3(gmin), _vec3(gmax)
        self.ctx.check(_L.mrmd_cell_list_build(self.h, atoms.h, begin, end, _p(d), _p(lo), _p(hi)))

    def __del__(self):
        if getattr(self, "h", None):
            _L.mrmd_cell_list_destroy(self.h)
            self.h = None

    def dims(self):
        d = np.zeros(3, dtype=np.int32)
        self.ctx.check(_L.mrmd_cell_list_dims(self.h, _p(d)))
        return d

    def download(self):
        nc = int(np.prod(self.dims()))
        n = self.end - self.begin
        cell = np.zeros(n, np.int32)
        counts = np.zeros(nc, np.int32)
        offsets = np.zeros(nc, np.int64)
        perm = np.zeros(n, np.int64)
        self.ctx.check(_L.mrmd_cell_list_download(self.h, _p(cell), _p(counts), _p(offsets), _p(perm)))
        return cell, counts, offsets, perm


class HalfVerletList:
    """HalfVerletList = Cabana::VerletList<HalfNeighborTag, VerletLayout2D, TeamOpTag>"""

    def __init__(self, ctx):
        self.ctx = ctx
        h = C.c_void_p()
        ctx.check(_L.mrmd_verlet_create(ctx.h, C.byref(h)))
        self.h = h

    def __del__(self):
        if getattr(self, "h", None):
            _L.mrmd_verlet_destroy(self.h)
            self.h = None

    def build(self, atoms, begin, end, r_n, cell_ratio, gmin, gmax, max_neigh=0, molecules=False):
        lo, hi = _vec3(gmin), _vec3(gmax)
        f = _L.mrmd_verlet_build_molecules if molecules else _L.mrmd_verlet_build
        self.ctx.check(f(self.h, atoms.h, begin, end, float(r_n), float(cell_ratio), _p(lo), _p(hi), max_neigh))

    def info(self):
        rows, mx, tot = C.c_int64(), C.c_int32(), C.c_int64()
        dims = np.zeros(3, np.int32)
        self.ctx.check(_L.mrmd_verlet_info(self.h, C.byref(rows), C.byref(mx), C.byref(tot), _p(dims)))
        return dict(rows=rows.value, max_full=mx.value, total_full=tot.value, dims=dims)

    def set_tiling(self, tile_cells=0, block=256, tile_columns=1):
        """cells per tile along z (0 = from the density), tile width in cell columns and
        workgroup size of the tiled list build and force kernel; applies from the next build"""
        self.ctx.check(_L.mrmd_verlet_set_tiling(self.h, int(tile_cells), int(tile_columns), int(block)))

    def set_stage_limit(self, limit=0):
        """most records a tile stages in LDS (0 = 2048); a tile that does not fit even at one
        cell column by one cell stages in global memory (dense inputs)"""
        self.ctx.check(_L.mrmd_verlet_set_stage_limit(self.h, int(limit)))

    def set_wide_offsets(self, on=True):
        """build with the 64-bit entry-offset form of the list kernel (a test knob: that form
        otherwise serves only lists of 2^31 entries or more); the list is the same"""
        self.ctx.check(_L.mrmd_verlet_set_wide_offsets(self.h, 1 if on else 0))

    def tiling(self):
        k, a, b, s = C.c_int(), C.c_int(), C.c_int(), C.c_int()
        self.ctx.check(_L.mrmd_verlet_get_tiling(self.h, C.byref(k), C.byref(a), C.byref(b), C.byref(s)))
        return dict(tile_cells=k.value, tile_columns=a.value, block=b.value, max_staged=s.value)

    def count_within(self, atoms, r):
        """(half-list pairs, both-direction entries) closer than r at the current positions"""
        h, f = C.c_int64(), C.c_int64()
        self.ctx.check(_L.mrmd_verlet_count_within(self.h, atoms.h, float(r), C.byref(h), C.byref(f)))
        return h.value, f.value

    def half_counts(self, n):
        out = np.zeros(n, np.int32)
        tot = C.c_int64()
        self.ctx.check(_L.mrmd_verlet_half_counts(self.h, _p(out), n, C.byref(tot)))
        return out, tot.value

    def num_pairs(self):
        tot = C.c_int64()
        self.ctx.check(_L.mrmd_verlet_half_counts(self.h, None, 0, C.byref(tot)))
        return tot.value

    def half_rows(self, n):
        counts, _ = self.half_counts(n)
        width = max(int(counts.max()) if n else 1, 1)
        out = np.full((n, width), -1, np.int32)
        self.ctx.check(_L.mrmd_verlet_half_neighbors(self.h, _p(out), n, width))
        return [set(out[i, :counts[i]].tolist()) for i in range(n)]

    def full_rows(self, n):
        info = self.info()
        width = max(info["max_full"], 1)
        out = np.full((n, width), -1, np.int32)
        counts = np.zeros(n, np.int32)
        self.ctx.check(_L.mrmd_verlet_full_neighbors(self.h, _p(out), _p(counts), n, width))
        return out, counts


def _vec(v, n):
    return np.ascontiguousarray(np.broadcast_to(np.asarray(v, dtype=np.float64), (n,)))


class LennardJones:
    """action::LennardJones (LennardJones.hpp:98-133); LennardJones(rc, sigma, eps, cap) is
    the unshifted single-type form, LennardJones.vector(...) the multi-type one"""

    def __init__(self, ctx, rc, sigma, epsilon, capping_distance=0.0, num_types=1, shifted=False,
                 type_stride=1):
        self.ctx = ctx
        nt2 = num_types * num_types
        self._arrs = [_vec(capping_distance, nt2), _vec(rc, nt2), _vec(sigma, nt2), _vec(epsilon, nt2)]
        h = C.c_void_p()
        ctx.check(_L.mrmd_lj_create(ctx.h, *[_p(a) for a in self._arrs], num_types, 1 if shifted else 0,
                                    type_stride, C.byref(h)))
        self.h = h
        self.num_types = num_types

    @classmethod
    def vector(cls, ctx, cap, rc, sigma, eps, num_types, shifted):
        # the reference's vector constructor keeps numTypes_ = 1 for the pair index
        return cls(ctx, rc, sigma, eps, cap, num_types, shifted, type_stride=1)

    def __del__(self):
        if getattr(self, "h", None):
            _L.mrmd_lj_destroy(self.h)
            self.h = None

    def apply(self, atoms, verlet_list, pred=None):
        self.ctx.check(_L.mrmd_lj_apply(self.h, atoms.h, verlet_list.h, C.byref(pred) if pred else None))

    def energy_virial(self):
        e, w = C.c_double(), C.c_double()
        self.ctx.check(_L.mrmd_lj_energy_virial(self.h, C.byref(e), C.byref(w)))
        return e.value, w.value

    def get_energy(self):
        return self.energy_virial()[0]

    def get_virial(self):
        return self.energy_virial()[1]

    def table(self):
        out = np.zeros((self.num_types ** 2, 10))
        self.ctx.check(_L.mrmd_lj_table(self.h, _p(out)))
        return out

    def eval(self, dist_sqr, type_idx=0):
        d = np.ascontiguousarray(np.atleast_1d(dist_sqr), dtype=np.float64)
        t = np.ascontiguousarray(np.broadcast_to(type_idx, d.shape), dtype=np.int64)
        ff, e = np.zeros_like(d), np.zeros_like(d)
        self.ctx.check(_L.mrmd_lj_eval(self.h, _p(d), _p(t), d.size, _p(ff), _p(e)))
        return ff, e

    def set_params(self, sigma, epsilon):
        s = _vec(sigma, self.num_types ** 2)
        e = _vec(epsilon, self.num_types ** 2)
        self.ctx.check(_L.mrmd_lj_set_params(self.h, _p(s), _p(e)))

    def adjoint(self, atoms, verlet_list, dLdF, ghost_layer=None):
        """host arrays: returns dL/dx (nlocal x 3), dL/deps, dL/dsigma (num_types^2)"""
        g = np.ascontiguousarray(dLdF, dtype=np.float64).reshape(-1, 3)
        dx = np.zeros_like(g)
        de = np.zeros(self.num_types ** 2)
        ds = np.zeros(self.num_types ** 2)
        self.ctx.check(_L.mrmd_lj_adjoint(self.h, atoms.h, verlet_list.h,
                                          ghost_layer.h if ghost_layer else None, _p(g), _p(dx), _p(de),
                                          _p(ds), 0))
        return dx, de, ds

    def adjoint_device(self, atoms, verlet_list, dLdF_ptr, dLdx_ptr, ghost_layer=None):
        de = np.zeros(self.num_types ** 2)
        ds = np.zeros(self.num_types ** 2)
        self.ctx.check(_L.mrmd_lj_adjoint(self.h, atoms.h, verlet_list.h,
                                          ghost_layer.h if ghost_layer else None, C.c_void_p(dLdF_ptr),
                                          C.c_void_p(dLdx_ptr), _p(de), _p(ds), 1))
        return de, ds


class VelocityVerlet:
    """action::VelocityVerlet (VelocityVerlet.hpp:26-31)"""

    @staticmethod
    def pre_force_integrate(atoms, dt):
        d = C.c_double()
        atoms.ctx.check(_L.mrmd_vv_pre(atoms.h, dt, C.byref(d)))
        return d.value

    @staticmethod
    def post_force_integrate(atoms, dt):
        atoms.ctx.check(_L.mrmd_vv_post(atoms.h, dt))


class VelocityVerletLangevinThermostat:
    """action::VelocityVerletLangevinThermostat(zeta, T); noise keyed by (seed, call, atom)"""

    def __init__(self, zeta, temperature, seed=1234):
        self.zeta, self.temperature, self.seed, self.call = zeta, temperature, seed, 0

    def set(self, zeta, temperature):
        self.zeta, self.temperature = zeta, temperature

    def pre_force_integrate(self, atoms, dt, pred=None):
        d = C.c_double()
        atoms.ctx.check(_L.mrmd_langevin_pre(atoms.h, dt, self.zeta, self.temperature, self.seed, self.call,
                                             C.byref(pred) if pred else None, C.byref(d)))
        self.call += 1
        return d.value

    def post_force_integrate(self, atoms, dt):
        VelocityVerlet.post_force_integrate(atoms, dt)


class LJ_IdealGas:
    """action::LJ_IdealGas (LJ_IdealGas.hpp:34-104)"""

    def __init__(self, ctx, capping_distance, rc, sigma, epsilon, do_shift=True, num_types=1):
        self.ctx = ctx
        nt2 = num_types * num_types
        self._arrs = [_vec(capping_distance, nt2), _vec(rc, nt2), _vec(sigma, nt2), _vec(epsilon, nt2)]
        h = C.c_void_p()
        ctx.check(_L.mrmd_ljig_create(ctx.h, *[_p(a) for a in self._arrs], num_types, 1 if do_shift else 0,
                                      C.byref(h)))
        self.h = h
        self.num_types = num_types

    def __del__(self):
        if getattr(self, "h", None):
            _L.mrmd_ljig_destroy(self.h)
            self.h = None

    def set_intervals(self, sampling, update):
        self.ctx.check(_L.mrmd_ljig_set_intervals(self.h, sampling, update))

    def set_queue(self, capacity):
        """tiled form: per-workgroup queue of pairs with a non-row partner (0..4096)"""
        self.ctx.check(_L.mrmd_ljig_set_queue(self.h, int(capacity)))

    def run(self, molecules, verlet_list, atoms):
        e = C.c_double()
        self.ctx.check(_L.mrmd_ljig_run(self.h, molecules.h, verlet_list.h, atoms.h, C.byref(e)))
        return e.value

    def histogram(self, which):
        out = np.zeros((200, self.num_types))
        self.ctx.check(_L.mrmd_ljig_get_histogram(self.h, which, _p(out)))
        return out

    def set_histogram(self, which, arr):
        a = np.ascontiguousarray(arr, dtype=np.float64)
        self.ctx.check(_L.mrmd_ljig_set_histogram(self.h, which, _p(a)))

    @property
    def run_counter(self):
        """run() calls so far (they decide the compensation sampling and update steps)"""
        n = C.c_int64()
        self.ctx.check(_L.mrmd_ljig_run_counter(self.h, C.byref(n), None))
        return n.value

    @run_counter.setter
    def run_counter(self, value):
        n = C.c_int64(int(value))
        self.ctx.check(_L.mrmd_ljig_run_counter(self.h, None, C.byref(n)))


class ThermodynamicForce:
    """action::ThermodynamicForce (ThermodynamicForce.hpp:32-96)"""

    def __init__(self, ctx, target_density, box, bin_width, modulation, enforce_symmetry=False,
                 use_periodicity=False):
        self.ctx = ctx
        t = np.ascontiguousarray(np.atleast_1d(target_density), dtype=np.float64)
        m = np.ascontiguousarray(np.atleast_1d(modulation), dtype=np.float64)
        self.num_types = t.size
        self.lo, self.hi = float(box.min[0]), float(box.max[0])  # the bins span x of the subdomain
        h = C.c_void_p()
        ctx.check(_L.mrmd_tf_create(ctx.h, _p(t), _p(m), t.size, C.byref(box), bin_width,
                                    int(enforce_symmetry), int(use_periodicity), C.byref(h)))
        self.h = h

    def __del__(self):
        if getattr(self, "h", None):
            _L.mrmd_tf_destroy(self.h)
            self.h = None

    @property
    def num_bins(self):
        n = C.c_int64()
        self.ctx.check(_L.mrmd_tf_num_bins(self.h, C.byref(n)))
        return n.value

    @property
    def samples(self):
        n = C.c_int64()
        self.ctx.check(_L.mrmd_tf_samples(self.h, C.byref(n)))
        return n.value

    def sample(self, atoms):
        self.ctx.check(_L.mrmd_tf_sample(self.h, atoms.h))

    def update(self, sigma, rng, pred=None):
        self.ctx.check(_L.mrmd_tf_update(self.h, sigma, rng, C.byref(pred) if pred else None))

    def apply(self, atoms, pred=None):
        self.ctx.check(_L.mrmd_tf_apply(self.h, atoms.h, C.byref(pred) if pred else None))

    def apply_interpolated(self, atoms, pred=None):
        self.ctx.check(_L.mrmd_tf_apply_interpolated(self.h, atoms.h, C.byref(pred) if pred else None))

    def _get(self, which):
        out = np.zeros((self.num_bins, self.num_types))
        self.ctx.check(_L.mrmd_tf_get(self.h, which, _p(out)))
        return out

    def get_force(self):
        return self._get(0)

    def get_density_profile(self):
        return self._get(1)

    def set_force(self, f):
        a = np.ascontiguousarray(f, dtype=np.float64)
        self.ctx.check(_L.mrmd_tf_set_force(self.h, _p(a)))

    def set_density_profile(self, density, samples):
        """restore the accumulated density profile and its sample count (checkpoints)"""
        a = np.ascontiguousarray(density, dtype=np.float64).reshape(self.num_bins, self.num_types)
        self.ctx.check(_L.mrmd_tf_set_density(self.h, _p(a), int(samples)))

    def grid(self):
        """bin centres, MultiHistogram::getBinPosition (MultiHistogram.hpp:68-74)"""
        nb = self.num_bins
        size = (self.hi - self.lo) / nb
        return self.lo + (np.arange(nb) + 0.5) * size

    def mu(self):
        l, r = np.zeros(self.num_types), np.zeros(self.num_types)
        self.ctx.check(_L.mrmd_tf_mu(self.h, _p(l), _p(r)))
        return l, r


def hist_gradient(ctx, lo, hi, data, periodic=False):
    d = np.ascontiguousarray(data, dtype=np.float64)
    out = np.zeros_like(d)
    ctx.check(_L.mrmd_hist_gradient(ctx.h, lo, hi, d.shape[0], d.shape[1], _p(d), int(periodic), _p(out)))
    return out


def hist_smoothen(ctx, lo, hi, data, sigma, rng, periodic=False):
    d = np.ascontiguousarray(data, dtype=np.float64)
    out = np.zeros_like(d)
    ctx.check(_L.mrmd_hist_smoothen(ctx.h, lo, hi, d.shape[0], d.shape[1], _p(d), sigma, rng, int(periodic),
                                    _p(out)))
    return out
