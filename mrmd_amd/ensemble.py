"""Ensemble engine: many independent periodic replicas advanced together (mrmd_batch_*).

`Ensemble` wraps the C ABI; `ensemble_config` draws the replicas of a thermodynamic-force
or parameter-sweep ensemble (atom counts log-uniform, densities and temperatures uniform,
half plain LJ NVT and half AdResS slabs with the proportions of the 200 x 40 x 40 slab:
atomistic diameter 0.15 Lx, hybrid width 0.05 Lx) with uniform-random initial positions.
"""
import ctypes as C

import numpy as np

from . import FORCE, POS, VEL, _L, _p
from .capi import Replica
from .lattice import maxwell_boltzmann, rng


_INT_FIELDS = (C.c_int32, C.c_int64, C.c_uint64)


def _replica_column(name, ctype, replicas):
    """one mrmd_replica field over all replicas, in a dtype that holds it exactly"""
    if name == "box":
        return np.array([[float(x) for x in r.get("box", (0.0, 0.0, 0.0))] for r in replicas])
    if ctype is C.c_uint64:
        return np.array([int(r.get(name, 0)) for r in replicas], np.uint64)
    if ctype in _INT_FIELDS:
        return np.array([int(r.get(name, 0)) for r in replicas], np.int64)
    return np.array([float(r.get(name, 0.0)) for r in replicas])


class Ensemble:
    def __init__(self, ctx, replicas, rc=2.5, skin=0.3):
        """replicas: sequence of dicts with the fields of mrmd_replica"""
        self.ctx = ctx
        self.n_rep = len(replicas)
        arr = (Replica * self.n_rep)()
        for k, r in enumerate(replicas):
            for name, _ in Replica._fields_:
                if name in r:
                    v = r[name]
                    if name == "box":
                        for d in range(3):
                            arr[k].box[d] = float(v[d])
                    else:
                        setattr(arr[k], name, v)
        self.replicas = [dict(r) for r in replicas]
        self.rc, self.skin = float(rc), float(skin)
        self.lj = dict(cap=0.0, sigma=1.0, epsilon=1.0, shifted=True)
        self.tiling_set = None
        self.stage_cap = 0
        self.counts = np.array([int(r["num_atoms"]) for r in replicas], dtype=np.int64)
        self.offsets = np.concatenate([[0], np.cumsum(self.counts)])
        self.n_total = int(self.offsets[-1])
        h = C.c_void_p()
        ctx.check(_L.mrmd_batch_create(ctx.h, arr, self.n_rep, float(rc), float(skin), C.byref(h)))
        self.h = h

    def __del__(self):
        if getattr(self, "h", None):
            _L.mrmd_batch_destroy(self.h)
            self.h = None

    def set_lj(self, cap=0.0, sigma=1.0, epsilon=1.0, shifted=True):
        self.ctx.check(_L.mrmd_batch_set_lj(self.h, float(cap), float(sigma), float(epsilon), 1 if shifted else 0))
        self.lj = dict(cap=float(cap), sigma=float(sigma), epsilon=float(epsilon), shifted=bool(shifted))

    def set_tiling(self, target_staged, block):
        self.ctx.check(_L.mrmd_batch_set_tiling(self.h, int(target_staged), int(block)))
        self.tiling_set = (int(target_staged), int(block))

    def set_stage_cap(self, limit):
        """cap on the LDS stage area in records (0 = automatic); larger tiles read global memory"""
        self.ctx.check(_L.mrmd_batch_set_stage_cap(self.h, int(limit)))
        self.stage_cap = int(limit)

    # ---- checkpoints ------------------------------------------------------------------
    def state(self):
        """the dynamic state as bytes (mrmd_batch_state_save); ends the current list epoch"""
        n = C.c_int64()
        self.ctx.check(_L.mrmd_batch_state_bytes(self.h, C.byref(n)))
        buf = np.zeros(n.value, np.uint8)
        self.ctx.check(_L.mrmd_batch_state_save(self.h, _p(buf), n.value))
        return buf

    def load_state(self, buf):
        b = np.ascontiguousarray(buf, dtype=np.uint8)
        self.ctx.check(_L.mrmd_batch_state_load(self.h, _p(b), b.size))

    def save(self, path):
        """checkpoint file (.npz): replica descriptions, cut-offs, LJ table and launch shape,
        and the dynamic state; Ensemble.restore(ctx, path) resumes from it bit-identically
        to this ensemble going on from here"""
        cols = {f"replica/{name}": _replica_column(name, t, self.replicas) for name, t in Replica._fields_}
        tiling = self.tiling_set if self.tiling_set is not None else (-1, -1)
        np.savez(path, state=self.state(), cutoffs=np.array([self.rc, self.skin]),
                 lj=np.array([self.lj["cap"], self.lj["sigma"], self.lj["epsilon"], float(self.lj["shifted"])]),
                 launch=np.array([tiling[0], tiling[1], self.stage_cap], np.int64), **cols)

    @classmethod
    def restore(cls, ctx, path):
        with np.load(path, allow_pickle=False) as f:
            cols = {name: f[f"replica/{name}"] for name, _ in Replica._fields_}
            reps = []
            for i in range(cols["num_atoms"].shape[0]):
                r = {}
                for name, t in Replica._fields_:
                    v = cols[name][i]
                    if name == "box":
                        r[name] = [float(x) for x in v]
                    elif t in _INT_FIELDS:
                        r[name] = int(v)
                    else:
                        r[name] = float(v)
                reps.append(r)
            rc, skin = (float(x) for x in f["cutoffs"])
            lj = f["lj"]
            launch = f["launch"]
            state = f["state"]
            e = cls(ctx, reps, rc=rc, skin=skin)
            e.set_lj(cap=float(lj[0]), sigma=float(lj[1]), epsilon=float(lj[2]), shifted=bool(lj[3]))
            if launch[0] >= 0:
                e.set_tiling(int(launch[0]), int(launch[1]))
            if launch[2]:
                e.set_stage_cap(int(launch[2]))
            e.load_state(state)
        return e

    def tiling(self):
        t, b = C.c_int(), C.c_int()
        k = np.zeros(self.n_rep, np.int32)
        self.ctx.check(_L.mrmd_batch_get_tiling(self.h, C.byref(t), C.byref(b), _p(k)))
        return dict(target_staged=t.value, block=b.value, tile_cells=k)

    def _concat(self, arrays):
        if isinstance(arrays, np.ndarray) and arrays.ndim == 2:
            out = arrays
        else:
            out = np.concatenate([np.asarray(a, dtype=np.float64).reshape(-1, 3) for a in arrays])
        out = np.ascontiguousarray(out, dtype=np.float64)
        assert out.shape == (self.n_total, 3)
        return out

    def upload(self, field, arrays):
        a = self._concat(arrays)
        self.ctx.check(_L.mrmd_batch_upload(self.h, field, _p(a)))

    def download(self, field):
        out = np.zeros((self.n_total, 3))
        self.ctx.check(_L.mrmd_batch_download(self.h, field, _p(out)))
        return out

    def split(self, a):
        return [a[self.offsets[k]:self.offsets[k + 1]] for k in range(self.n_rep)]

    def positions(self):
        return self.split(self.download(POS))

    def velocities(self):
        return self.split(self.download(VEL))

    def forces(self):
        return self.split(self.download(FORCE))

    def run(self, nsteps, graph_steps=10):
        self.ctx.check(_L.mrmd_batch_run(self.h, int(nsteps), int(graph_steps)))

    def energies(self):
        pe, ke = np.zeros(self.n_rep), np.zeros(self.n_rep)
        self.ctx.check(_L.mrmd_batch_energies(self.h, _p(pe), _p(ke)))
        return pe, ke

    def stats(self):
        reb = np.zeros(self.n_rep, np.int64)
        steps = C.c_int64()
        self.ctx.check(_L.mrmd_batch_stats(self.h, _p(reb), C.byref(steps)))
        return reb, steps.value

    def tf(self, replica, which=0):
        nb = int(np.ceil(self.replicas[replica]["box"][0] / self.replicas[replica]["tf_bin_width"]))
        out = np.zeros(nb)
        self.ctx.check(_L.mrmd_batch_tf_get(self.h, int(replica), int(which), _p(out), nb))
        return out


def adress_box(n, rho, rn=2.8, aspect=5.0):
    """edges of an AdResS slab box of n atoms at density rho with Lx : Ly = aspect where the
    short edges allow it (every edge at least 2 (rc + skin))"""
    vol = n / rho
    ly = max((vol / aspect) ** (1.0 / 3.0), 2.0 * rn + 1e-6)
    lx = max(vol / (ly * ly), 2.0 * rn + 1e-6)
    return [lx, ly, ly]


def cubic_box(n, rho, rn=2.8):
    edge = max((n / rho) ** (1.0 / 3.0), 2.0 * rn + 1e-6)
    return [edge] * 3


def replica(n, box, temperature, seed, adress=False, dt=0.005, gamma=1.0, tf_every=100):
    r = dict(num_atoms=int(n), box=list(box), dt=dt, gamma=gamma, temperature=temperature, mass=1.0,
             seed=int(seed), adress=1 if adress else 0, tf_every=tf_every if adress else 0,
             atomistic_diameter=0.15 * box[0], hybrid_diameter=0.05 * box[0], nu=1, tf_bin_width=1.0,
             tf_target_density=n / float(np.prod(box)), tf_sigma=2.0, tf_range=3.0, tf_modulation=1.0,
             tf_symmetric=1, tf_periodic=1)
    return r


def relax(ctx, reps, pos, steps=200, cap=0.7, gamma=50.0, seed=11, rc=2.5, skin=0.3):
    """relax uniform-random starts with the capped potential under strong Langevin friction
    (the overlap energy leaves within a few steps), then draw fresh Maxwell-Boltzmann
    velocities at each replica's temperature; returns (positions, velocities)"""
    strong = [dict(r, gamma=gamma) for r in reps]
    e = Ensemble(ctx, strong, rc, skin)
    e.set_lj(cap=cap)
    e.upload(POS, pos)
    e.upload(VEL, [np.zeros((r["num_atoms"], 3)) for r in reps])
    e.run(steps, 10)
    p = e.positions()
    del e
    vel = [maxwell_boltzmann(r["num_atoms"], r["temperature"], seed * 7919 + k) for k, r in enumerate(reps)]
    return p, vel


def ensemble_config(n_rep=256, n_min=500, n_max=50000, rho=(0.3, 0.9), temp=(0.7, 1.5), seed=2024,
                    adress_fraction=0.5, rn=2.8):
    """replica dicts, initial positions (uniform random in the box) and Maxwell-Boltzmann
    velocities, all drawn from the seeded generator"""
    g = rng(seed)
    reps, pos, vel = [], [], []
    for k in range(n_rep):
        n = int(round(np.exp(g.uniform(np.log(n_min), np.log(n_max)))))
        d = g.uniform(*rho)
        t = g.uniform(*temp)
        ad = k < int(round(adress_fraction * n_rep))
        box = adress_box(n, d, rn) if ad else cubic_box(n, d, rn)
        reps.append(replica(n, box, t, seed * 1000 + k, adress=ad))
        pos.append(g.uniform(0.0, 1.0, (n, 3)) * np.asarray(box))
        vel.append(maxwell_boltzmann(n, t, seed * 7919 + k))
    return reps, pos, vel
