"""mrmd_lj_apply_noalloc (the allocation-free LennardJones::apply) through the C ABI against
mrmd_lj_apply on the same atoms and list: forces, energy and virial bit-identical, on two
calls in a row, the second on a smaller system (the shared scratch array keeps the first
call's partials, which the second must not see).  Energy and virial land in device memory
the caller owns (allocated here through the HIP runtime the library links)."""
import ctypes as C

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


class DeviceDoubles:
    """n doubles of device memory from the HIP runtime the library itself links"""

    def __init__(self, n):
        self.hip = C.CDLL("libamdhip64.so.7")
        self.n = n
        self.p = C.c_void_p()
        assert self.hip.hipMalloc(C.byref(self.p), C.c_size_t(8 * n)) == 0

    def read(self):
        out = np.zeros(self.n)
        assert self.hip.hipMemcpy(out.ctypes.data_as(C.c_void_p), self.p, C.c_size_t(8 * self.n), 2) == 0
        return out

    def free(self):
        self.hip.hipFree(self.p)


def _system(M, ctx, cells, seed, cap=0.0):
    from mrmd_amd.lattice import fcc
    pos, L = fcc(cells, 0.8442)
    pos = np.mod(pos + np.random.default_rng(seed).uniform(-0.15, 0.15, pos.shape), L)
    a = M.Atoms.from_arrays(ctx, pos)
    box = M.make_box([0, 0, 0], [L] * 3, 2.8)
    gl = M.GhostLayer(ctx)
    gl.create_ghost_atoms(a, box)
    vl = M.HalfVerletList(ctx)
    vl.build(a, 0, a.nlocal, 2.8, 1.0, [-2.8] * 3, [L + 2.8] * 3)
    lj = M.LennardJones(ctx, 2.5, 1.0, 1.0, capping_distance=cap, shifted=True)
    return a, gl, vl, lj


def _compare(M, ctx, a, vl, lj, ev_dev):
    a.set_force(0.0)
    lj.apply(a, vl)
    f_ref = a.download(M.FORCE, a.nlocal)
    e_ref, w_ref = lj.energy_virial()
    a.set_force(0.0)
    rc = M._L.mrmd_lj_apply_noalloc(lj.h, a.h, vl.h, ev_dev.p)
    assert rc == 0, M._L.mrmd_ctx_last_error(ctx.h)
    ctx.synchronize()
    ev = ev_dev.read()
    assert np.array_equal(a.download(M.FORCE, a.nlocal), f_ref)
    assert ev[0] == e_ref and ev[1] == w_ref, (ev, e_ref, w_ref)
    # the table's own energy / virial follow the allocation-free call too
    assert lj.energy_virial() == (e_ref, w_ref)


def test_noalloc_apply_equals_apply_twice_smaller_second(M, ctx):
    ev_dev = DeviceDoubles(2)
    try:
        big = _system(M, ctx, 20, 1)            # 32,000 atoms
        small = _system(M, ctx, 8, 2, cap=0.8)  # 2,048 atoms, capped table (the other kernel form)
        assert big[2].info()["rows"] > small[2].info()["rows"]
        _compare(M, ctx, big[0], big[2], big[3], ev_dev)
        _compare(M, ctx, small[0], small[2], small[3], ev_dev)
    finally:
        ctx.synchronize()
        ev_dev.free()


def test_noalloc_apply_refuses_asymmetric_tables(M, ctx):
    ev_dev = DeviceDoubles(2)
    a, gl, vl, _ = _system(M, ctx, 6, 3)
    # a two-type table with table(0, 1) != table(1, 0) (index t_i * 2 + t_j)
    h = C.c_void_p()
    cap = (C.c_double * 4)(0, 0, 0, 0)
    rc = (C.c_double * 4)(2.5, 2.5, 2.5, 2.5)
    sig = (C.c_double * 4)(1.0, 0.9, 0.8, 1.1)
    eps = (C.c_double * 4)(1.0, 0.5, 0.7, 0.6)
    assert M._L.mrmd_lj_create(ctx.h, cap, rc, sig, eps, 2, 1, 2, C.byref(h)) == 0
    try:
        assert M._L.mrmd_lj_apply_noalloc(h, a.h, vl.h, ev_dev.p) == -5
    finally:
        M._L.mrmd_lj_destroy(h)
        ev_dev.free()
