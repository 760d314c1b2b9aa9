"""Throughput benchmark of the plain LJ hot path on one MI355X.

Workload (the reference hot path's headline throughput configuration): N = 4,000,000 atoms
on an FCC lattice (100^3 cells x 4) at density 0.8442, FP64, Lennard-Jones rc = 2.5 sigma
(energy-shifted), Verlet skin 0.3, Langevin thermostat T = 1.0, gamma = 1.0, dt = 0.005.
One step = Langevin pre-force integration, skin check (with wrap + cell permutation +
ordered ghost creation + Verlet rebuild when due, otherwise a ghost refresh), force
zeroing, LJ force/energy/virial, ghost force fold, post-force integration: the loop of the
reference's LJ examples, through the library's action API (the timed steps run it from
C++, mrmd_md_run, the calls a C++ caller of the actions makes; the rebuild decision reads
the displacement back every step as there).  HIP events surround only the force kernel
during the timed steps; the other kernels' times come from an untimed pass afterwards.

Prints one JSON line.  `value` is atom-steps/s of the timed steps with all data resident
in HBM.  The launch shape of the tiled list and force kernels is autotuned during warm-up.  `roofline` reports the LJ force kernel (algorithmic FP64 FLOPs of the reference's
half-list algorithm, 8 per candidate pair + 24 per pair inside rc with energy and virial,
divided by the kernel's mean duration from HIP events on the library stream).
`cpu_baseline` times the CPU restatement of the reference algorithm (AoS, half list with
atomic F_j, OpenMP) on a bounded sample of the same system.  `batched_vs_serial_speedup`
compares the ensemble engine on a 256-replica ensemble with the same replicas run one after
another through the single-system path; `adjoint_over_forward` is the adjoint LJ kernel's
time over the forward force kernel's on a 32,000-atom two-type mixture.  `extras.adress_slab`
times the AdResS LJ / ideal-gas slab (160,000 molecules in a 200 x 40 x 40 box) through the
single-system path, beside the CPU restatement on a bounded sample of the same run
(`--adress-only` runs just that workload).
"""
import argparse
import json
import os
import subprocess
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

import numpy as np  # noqa: E402

FP64_PEAK_TFLOPS = 78.6
HBM_PEAK_GBS = 8000.0


def measured_traffic(key="traffic_bytes_per_launch"):
    """HBM bytes per launch from the committed rocprofv3 FETCH_SIZE / WRITE_SIZE summary of
    this workload (calibrated against a device copy): `key` names the force kernel's line
    (default) or the list build's, or None"""
    path = os.path.join(ROOT, "profiles", "lj_force_4M_traffic.md")
    try:
        with open(path) as f:
            for line in f:
                if line.startswith(key + ":"):
                    return float(line.split(":", 1)[1])
    except OSError:
        return None
    return None


def build_system(M, ctx, cells, temperature, seed):
    from mrmd_amd.drivers import LJSystem
    from mrmd_amd.lattice import fcc, maxwell_boltzmann
    pos, L = fcc(cells, 0.8442)
    vel = maxwell_boltzmann(pos.shape[0], temperature, seed)
    s = LJSystem(ctx, pos, vel, L, rc=2.5, skin=0.3, shifted=True, langevin=True, gamma=1.0,
                 temperature=temperature, seed=seed, permute=True)
    return s, pos, vel, L


def cpu_baseline(pos, vel, L, steps, warm, reps, threads):
    """the CPU restatement of the reference algorithm (AoS records, half list with atomic F_j,
    OpenMP) on the same system: `warm` untimed steps, then `reps` timed runs of `steps` steps
    each; returns [(atom-steps/s, seconds, rebuilds)] per run"""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import pyoracle as O
    O.lib().or_set_threads(threads)
    a = O.Atoms.from_arrays(pos, vel=vel)
    box = O.make_box([0, 0, 0], [L] * 3, 2.8)
    prm = O.md_params(nsteps=0, langevin=1, permute=1, shifted=1, gamma=1.0, temperature=1.0, seed=1234)
    md = O.lib().or_md_new(a.h, O.C.byref(box), O.C.byref(prm))
    O.lib().or_md_steps(md, warm)
    runs = []
    for _ in range(reps):
        r0 = O.lib().or_md_steps(md, 0)
        t0 = time.perf_counter()
        r1 = O.lib().or_md_steps(md, steps)
        dt = time.perf_counter() - t0
        runs.append((pos.shape[0] * steps / dt, dt, r1 - r0))
    O.lib().or_md_free(md)
    return runs


def host_cpu():
    """the host CPU: model and physical cores from lscpu, and the CPUs this process may use"""
    info = {"model": None, "physical_cores": None, "affinity_cpus": len(os.sched_getaffinity(0))}
    try:
        out = subprocess.run(["lscpu"], capture_output=True, text=True, timeout=30).stdout
        kv = {k.strip(): v.strip() for k, v in (l.split(":", 1) for l in out.splitlines() if ":" in l)}
        info["model"] = kv.get("Model name")
        info["physical_cores"] = int(kv["Core(s) per socket"]) * int(kv["Socket(s)"])
    except (OSError, KeyError, ValueError, subprocess.SubprocessError):
        pass
    return info


def cpu_threads():
    """(host_cpu(), the CPU share OMP_NUM_THREADS grants, threads to use): the physical cores,
    capped by the share and by this process's affinity"""
    cpu = host_cpu()
    share = int(os.environ.get("OMP_NUM_THREADS", cpu["affinity_cpus"]))
    threads = max(1, min(x for x in (cpu["physical_cores"], cpu["affinity_cpus"], share) if x))
    return cpu, share, threads


def cpu_baseline_leg(cells, steps, warm, reps):
    """runs the CPU sample in a child process, so that the OpenMP placement (threads spread
    over the cores, one per place) applies from the runtime's start: threads = the physical
    cores lscpu reports, capped by this process's CPU share (OMP_NUM_THREADS when set, and
    its affinity).  Returns the cpu_baseline object (median of the timed runs)."""
    cpu, share, threads = cpu_threads()
    env = dict(os.environ, OMP_NUM_THREADS=str(threads), OMP_PROC_BIND="spread", OMP_PLACES="threads")
    out = subprocess.run([sys.executable, os.path.abspath(__file__), "--cpu-sample", "--cells", str(cells),
                          "--cpu-steps", str(steps), "--cpu-warmup", str(warm), "--cpu-reps", str(reps),
                          "--cpu-threads", str(threads)], capture_output=True, text=True, env=env, check=True)
    runs = json.loads(out.stdout.strip().splitlines()[-1])["runs"]
    rates = sorted(r[0] for r in runs)
    med = rates[len(rates) // 2]
    n = 4 * cells ** 3
    return {"value": med, "unit": "atom-steps/s", "cores": threads, "kind": "port",
            "sample": f"median of {reps} runs of {steps} steps ({reps * steps} timed steps) of the same {n}-atom "
                      f"system after {warm} warm-up steps; {sum(r[2] for r in runs)} list rebuild(s) in the timed "
                      f"steps, {sum(r[1] for r in runs):.1f} s timed; OpenMP {threads} threads, "
                      f"OMP_PROC_BIND=spread, OMP_PLACES=threads",
            "runs_atom_steps_per_s": [r[0] for r in runs], "cpu_model": cpu["model"],
            "physical_cores_lscpu": cpu["physical_cores"], "process_cpus": cpu["affinity_cpus"],
            "cpu_share_omp_num_threads": share}


def ensemble_bench(ctx, n_rep, steps, tune=True):
    """256-replica ensemble (atom counts log-uniform in [500, 50000], densities in [0.3, 0.9],
    temperatures in [0.7, 1.5], half AdResS slabs) through the ensemble engine with 10-step
    graphs, against the same replicas run one after another through the single-system path
    (LJSystem / AdResSSystem) over the same number of steps from the same state.  Both use
    the capped potential (r_cap = 0.7) that keeps ideal-gas particles entering the hybrid
    region finite.  The launch shape comes from mrmd_amd.autotune.tune_ensemble."""
    from mrmd_amd.autotune import tune_ensemble
    from mrmd_amd.drivers import AdResSSystem, LJSystem
    from mrmd_amd.ensemble import Ensemble, ensemble_config, relax
    reps, pos, _ = ensemble_config(n_rep=n_rep, seed=2024)
    pos, vel = relax(ctx, reps, pos)  # capped relaxation of the uniform-random starts
    e = Ensemble(ctx, reps)
    e.set_lj(cap=0.7)
    e.upload(0, pos)
    e.upload(1, vel)
    e.run(100, 10)
    tuning = tune_ensemble(e)[0] if tune else None
    e.run(10, 10)
    ctx.synchronize()
    P, V = e.positions(), e.velocities()  # the serial runs start from this state too
    t0 = time.perf_counter()
    e.run(steps, 10)
    ctx.synchronize()
    batched_s = time.perf_counter() - t0
    n_total = e.n_total
    batched = n_total * steps / batched_s
    # serial: every replica through the single-system path, the same steps
    serial_time, serial_atom_steps = 0.0, 0
    for k, r in enumerate(reps):
        common = dict(rc=2.5, skin=0.3, cap=0.7, shifted=True, gamma=r["gamma"], temperature=r["temperature"],
                      seed=r["seed"])
        if r["adress"]:
            sysk = AdResSSystem(ctx, P[k], V[k], r["box"], r["atomistic_diameter"], r["hybrid_diameter"], r["nu"],
                                **common)
        else:
            sysk = LJSystem(ctx, P[k], V[k], r["box"], langevin=True, **common)
        sysk.run(1, r["dt"])  # its first list build, outside the clock like the ensemble's
        ctx.synchronize()
        t0 = time.perf_counter()
        sysk.run(steps, r["dt"])
        ctx.synchronize()
        serial_time += time.perf_counter() - t0
        serial_atom_steps += r["num_atoms"] * steps
        del sysk
    serial = serial_atom_steps / serial_time
    return {"replicas": n_rep, "atoms": n_total, "adress_replicas": sum(r["adress"] for r in reps),
            "batched_atom_steps_per_s": batched, "batched_steps": steps, "graph_steps": 10,
            "serial_atom_steps_per_s": serial, "serial_steps_per_replica": steps,
            "serial_note": "each replica constructed from the ensemble's state before its timed steps and run "
                           "one untimed step (its first list build, as the ensemble's lists and graphs exist before "
                           "its clock starts), then the same number of steps timed",
            "speedup": batched / serial, "tiling": tuning}


def adjoint_ratio(ctx, reps=5):
    """adjoint / forward kernel time on a 32,000-atom two-type LJ mixture (80/20, density
    0.8, Lorentz-Berthelot, rc 2.5), HIP events on the library stream"""
    from mrmd_amd.adjoint import Snapshot
    from mrmd_amd.lattice import fcc, rng
    pos, L = fcc(20, 0.8)
    g = rng(5)
    pos = np.mod(pos + g.uniform(-0.1, 0.1, pos.shape), L)
    types = (g.random(pos.shape[0]) < 0.2).astype(np.int64)
    eAB, sAB = np.sqrt(0.5), 0.5 * (1.0 + 0.88)
    snap = Snapshot(ctx, pos, types, L, 2, np.array([1.0, eAB, eAB, 0.5]), np.array([1.0, sAB, sAB, 0.88]))
    up = g.standard_normal((pos.shape[0], 3))
    snap.forces()
    snap.adjoint(up)
    ctx.set_profiling(True)
    ctx.reset_stats()
    for _ in range(reps):
        snap.forces()
        snap.adjoint(up)
    nf, tf = ctx.kernel_stats("lj_force")
    na, ta = ctx.kernel_stats("lj_adjoint")
    ctx.set_profiling(False)
    # whole calls as a caller sees them: forces downloaded to the host, the adjoint from a
    # host upstream gradient to host gradients (dL/dx and dL/d(eps, sigma))
    ctx.synchronize()
    t0 = time.perf_counter()
    for _ in range(reps):
        snap.forces()
    call_f = (time.perf_counter() - t0) / reps
    t0 = time.perf_counter()
    for _ in range(reps):
        snap.adjoint(up)
    call_a = (time.perf_counter() - t0) / reps
    return {"atoms": int(pos.shape[0]), "forward_ms": tf / nf, "adjoint_ms": ta / na, "ratio": (ta / na) / (tf / nf),
            "forward_call_ms": call_f * 1e3, "adjoint_call_ms": call_a * 1e3, "call_ratio": call_a / call_f}


ADRESS_L = (200.0, 40.0, 40.0)
ADRESS_RHO = 0.5
ADRESS_AT, ADRESS_HY = 30.0, 10.0  # atomistic diameter |x - 100| < 15, hybrid width 10
ADRESS_CAP = 0.7                   # capping radius of the reference's AdResS example


def adress_state(ctx):
    """the AdResS slab's start: 160,000 monatomic molecules uniform-random in the periodic
    200 x 40 x 40 box (density 0.5, seeded), relaxed 200 steps with the capped force"""
    from mrmd_amd.ensemble import relax, replica
    from mrmd_amd.lattice import rng
    L = np.asarray(ADRESS_L)
    n = int(round(ADRESS_RHO * float(np.prod(L))))
    pos = rng(7).uniform(0.0, 1.0, (n, 3)) * L
    rep = replica(n, list(L), 1.0, 7, adress=True, tf_every=100)
    rep.update(atomistic_diameter=ADRESS_AT, hybrid_diameter=ADRESS_HY)
    P, V = relax(ctx, [rep], [pos])
    return P[0], V[0], list(L)


def adress_slab(ctx, steps, warm=50, profile_steps=100):
    """AdResS LJ / ideal-gas slab through the single-system path (AdResSSystem: multi-
    resolution ghosts, molecule list, LJ_IdealGas, thermodynamic force on 1-sigma x-bins
    updated every 100 steps over the hybrid region, Langevin T = 1, gamma = 1, dt = 0.005).
    Timed steps run without per-kernel events; a separate pass of profile_steps steps gives
    the kernel breakdown."""
    from mrmd_amd.drivers import AdResSSystem
    P, V, L = adress_state(ctx)
    n = P.shape[0]
    s = AdResSSystem(ctx, P, V, L, ADRESS_AT, ADRESS_HY, 1, cap=ADRESS_CAP, seed=7)
    s.run(warm, 0.005)
    ctx.synchronize()
    r0 = s.rebuilds
    t0 = time.perf_counter()
    s.run(steps, 0.005)
    ctx.synchronize()
    elapsed = time.perf_counter() - t0
    rebuilds = s.rebuilds - r0
    names = ("lj_idealgas_tiled", "lj_idealgas", "verlet_build", "post_pre", "langevin_pre", "vv_post", "halo_update",
             "halo_fold", "update_molecules", "contribute_molecule_force", "tf_sample", "tf_apply", "adress_prepare",
             "adress_collect")
    ctx.set_profiling(True)
    ctx.reset_stats()
    s.run(profile_steps, 0.005)
    ctx.synchronize()
    kernels = {k: {"launches": ctx.kernel_stats(k)[0], "total_ms": ctx.kernel_stats(k)[1]} for k in names}
    ctx.set_profiling(False)
    return {"atoms": n, "box": L, "steps": steps, "atom_steps_per_s": n * steps / elapsed,
            "ms_per_step": elapsed / steps * 1e3, "rebuilds_in_timed_steps": rebuilds, "cap": ADRESS_CAP,
            "tiling": s.vl.tiling(), "kernel_ms_per_100_steps": kernels}, (P, V, L)


def cpu_adress_sample(P, V, L, steps, threads):
    """the same AdResS step loop through the CPU restatement (OpenMP), started from the
    same relaxed state: returns (atom-steps/s, seconds, rebuilds)"""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import pyoracle as O
    O.lib().or_set_threads(threads)
    run = O.AdResSRun(P, V, L, ADRESS_AT, ADRESS_HY, 1, cap=ADRESS_CAP, seed=7)
    t0 = time.perf_counter()
    for _ in range(steps):
        run.step(0.005)
    seconds = time.perf_counter() - t0
    return P.shape[0] * steps / seconds, seconds, run.rebuilds


def copy_calibration(mib):
    """a plain device-to-device copy of known size, for scaling rocprofv3's FETCH_SIZE /
    WRITE_SIZE counters to bytes"""
    import torch
    n = mib * 1024 * 1024 // 8
    src = torch.ones(n, dtype=torch.float64, device="cuda")
    dst = torch.empty_like(src)
    for _ in range(3):
        dst.copy_(src)
    torch.cuda.synchronize()
    print(json.dumps({"copy_calibration_bytes": n * 8, "copies": 3}))


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--cells", type=int, default=100, help="FCC cells per edge (100 -> 4M atoms)")
    ap.add_argument("--profile-steps", type=int, default=60,
                    help="untimed steps after the timed ones that time every kernel (HIP events)")
    ap.add_argument("--cpu-steps", type=int, default=17, help="timed steps per CPU run")
    ap.add_argument("--cpu-warmup", type=int, default=10)
    ap.add_argument("--cpu-reps", type=int, default=3, help="timed CPU runs (the median is reported)")
    ap.add_argument("--cpu-sample", action="store_true", help=argparse.SUPPRESS)
    ap.add_argument("--cpu-threads", type=int, default=0, help=argparse.SUPPRESS)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-tune", action="store_true", help="keep the default tiling (no autotuning)")
    ap.add_argument("--no-extras", action="store_true", help="skip the ensemble and adjoint figures")
    ap.add_argument("--ensemble-replicas", type=int, default=256)
    ap.add_argument("--ensemble-steps", type=int, default=300, help="steps of the batched and of each serial run")
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="MIB",
                    help="only run a device copy of MIB MiB (counter calibration) and exit")
    ap.add_argument("--adress-steps", type=int, default=300, help="timed steps of the AdResS slab extra")
    ap.add_argument("--adress-cpu-steps", type=int, default=20, help="CPU sample of the AdResS slab")
    ap.add_argument("--adress-only", action="store_true",
                    help="only run the AdResS slab (no CPU sample) and print its JSON line")
    ap.add_argument("--ensemble-only", action="store_true",
                    help="only run the ensemble comparison and print its JSON line")
    args = ap.parse_args()
    if args.cpu_sample:
        # the CPU leg's child process: the same FCC system and velocities, from the lattice
        # module alone (the package would load the GPU library)
        import importlib.util
        spec = importlib.util.spec_from_file_location("lattice", os.path.join(ROOT, "mrmd_amd", "lattice.py"))
        lattice = importlib.util.module_from_spec(spec)
        spec.loader.exec_module(lattice)
        pos, L = lattice.fcc(args.cells, 0.8442)
        vel = lattice.maxwell_boltzmann(pos.shape[0], 1.0, 1234)
        runs = cpu_baseline(pos, vel, L, args.cpu_steps, args.cpu_warmup, args.cpu_reps, args.cpu_threads)
        print(json.dumps({"runs": runs}))
        return
    if args.copy_calibration:
        copy_calibration(args.copy_calibration)
        return
    if args.adress_only:
        import mrmd_amd as M
        res, _ = adress_slab(M.Context(0), args.adress_steps)
        print(json.dumps({"adress_slab": res}))
        return
    if args.ensemble_only:
        import mrmd_amd as M
        res = ensemble_bench(M.Context(0), args.ensemble_replicas, args.ensemble_steps, tune=not args.no_tune)
        print(json.dumps({"ensemble": res}))
        return
    if args.gpus != 1:
        print(json.dumps({"error": "bench.py measures one GPU (the kernels on one GPU's share); runs on several GPUs are not implemented",
                          "n_gpus": args.gpus}))
        sys.exit(2)

    import mrmd_amd as M
    ctx = M.Context(0)
    s, pos, vel, L = build_system(M, ctx, args.cells, 1.0, 1234)
    n = pos.shape[0]
    dt = 0.005
    s.run(args.warmup, dt)
    tiling = None
    if not args.no_tune:
        from mrmd_amd.autotune import tune_lj_system
        best, _ = tune_lj_system(s)
        tiling = {"tile_columns": best[0], "tile_cells": best[1], "block": best[2]}
        s.run(3, dt)
    else:
        tiling = {k: v for k, v in s.vl.tiling().items() if k != "max_staged"}
    ctx.synchronize()
    # the timed steps carry HIP events around the force kernel only (the roofline's
    # measurement, on the library stream); the other kernels are timed in a separate pass
    ctx.set_profiling(True, only="lj_force")
    ctx.reset_stats()
    reb0 = s.rebuilds
    t0 = time.perf_counter()
    s.run(args.steps, dt)  # the example's time loop, natively (mrmd_md_run)
    ctx.synchronize()
    elapsed = time.perf_counter() - t0
    rebuilds = s.rebuilds - reb0
    launches, force_ms = ctx.kernel_stats("lj_force")
    ctx.set_profiling(True)
    ctx.reset_stats()
    s.run(args.profile_steps, dt)
    stats = {k: ctx.kernel_stats(k) for k in ("verlet_build", "verlet_tiled", "post_pre", "langevin_pre", "vv_post", "halo_create",
                                              "halo_update", "halo_fold", "atoms_permute", "cell_list_build")}
    nrec = s.atoms.nlocal + s.atoms.nghost
    full_entries = s.vl.info()["total_full"]
    ctx.set_profiling(False)
    half_rn, _ = s.vl.count_within(s.atoms, 2.8)
    half_rc, _ = s.vl.count_within(s.atoms, 2.5)
    flops = 8.0 * half_rn + 24.0 * half_rc
    avg_s = force_ms / max(launches, 1) / 1e3
    achieved_tf = flops / avg_s / 1e12
    value = n * args.steps / elapsed

    def gbs(name, nbytes):
        k, t = stats[name]
        return {"GB/s": nbytes / (t / k / 1e3) / 1e9 if k else None, "bytes": nbytes,
                "avg_ms": t / k if k else None, "frac_of_hbm_peak": nbytes / (t / k / 1e3) / 1e9 / HBM_PEAK_GBS
                if k else None}

    # algorithmic bytes: Langevin pre-force step 136 B/atom (read position 32, velocity 24,
    # force 24; write position 32, velocity 24), post-force kick 80 B/atom (read velocity,
    # force, mass; write velocity), and the fused post + next pre pass of the time loop the
    # same 136 B/atom as the pre-force step alone; list build: read 32 B + write 4 B cell
    # index + 4 B slot + 32 B staged copy per record, 2 B per list entry and 4 B per row count
    bandwidth = {"post_pre": gbs("post_pre", 136.0 * n), "langevin_pre": gbs("langevin_pre", 136.0 * n),
                 "vv_post": gbs("vv_post", 80.0 * n),
                 "verlet_build": gbs("verlet_build", 72.0 * nrec + 2.0 * full_entries + 4.0 * nrec)}
    # the tiled list kernel alone: each record's position read once (32 B), a row's slot ->
    # record and record -> cell lookups (8 B), 2 B per list entry and a 4 B count per row
    list_bytes = 32.0 * nrec + 12.0 * n + 2.0 * full_entries
    lk, lt = stats["verlet_tiled"]
    list_gbs = list_bytes / (lt / lk / 1e3) / 1e9 if lk else None
    roofline_list = {"bound": "hbm", "achieved": list_gbs, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": list_gbs / HBM_PEAK_GBS if lk else None,
                     "traffic": measured_traffic("list_build_traffic_bytes_per_launch") if args.cells == 100 else None,
                     "kernel": "k_verletTiled (Verlet list build)", "avg_ms": lt / lk if lk else None,
                     "launches": lk, "bytes_per_launch": list_bytes}

    cpu = None
    if not args.no_cpu:
        cpu = cpu_baseline_leg(args.cells, args.cpu_steps, args.cpu_warmup, args.cpu_reps)

    extras = None
    if not args.no_extras:
        extras = {"ensemble": ensemble_bench(ctx, args.ensemble_replicas, args.ensemble_steps, tune=not args.no_tune),
                  "adjoint": adjoint_ratio(ctx)}
        ad, (P, V, L2) = adress_slab(ctx, args.adress_steps)
        if not args.no_cpu:
            threads = cpu_threads()[2]
            cv, ct, creb = cpu_adress_sample(P, V, L2, args.adress_cpu_steps, threads)
            ad["cpu_baseline"] = {"value": cv, "unit": "atom-steps/s", "cores": threads, "kind": "port",
                                  "sample": f"{args.adress_cpu_steps} steps from the same relaxed state, "
                                            f"{creb} list rebuild(s) inside, {ct:.1f} s wall"}
        extras["adress_slab"] = ad

    out = {
        "metric": "atom-steps/s at 4M-atom LJ",
        "value": value,
        "unit": "atom-steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": elapsed / args.steps * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic FCC lattice + Maxwell-Boltzmann velocities (seeded)",
        "config": {"workload": "plain LJ NVT (Langevin), FP64, FCC 100^3 x 4", "atoms": n,
                   "density": 0.8442, "rc": 2.5, "skin": 0.3, "dt": dt, "temperature": 1.0, "gamma": 1.0,
                   "rebuilds_in_timed_steps": rebuilds, "tiling": tiling},
        "roofline": {"bound": "fp64_valu", "achieved": achieved_tf, "peak": FP64_PEAK_TFLOPS, "unit": "TFLOP/s",
                     "frac": achieved_tf / FP64_PEAK_TFLOPS,
                     "traffic": measured_traffic() if args.cells == 100 else None,
                     "kernel": "k_ljTiled (LJ force)", "avg_ms": avg_s * 1e3, "launches": launches,
                     "half_pairs_rn": half_rn, "half_pairs_rc": half_rc, "flop_per_launch": flops},
        "roofline_list_build": roofline_list,
        "cpu_baseline": cpu,
        "kernel_ms": {k: {"launches": v[0], "total_ms": v[1]} for k, v in stats.items()},
        "kernel_ms_pass": f"{args.profile_steps} untimed steps after the timed ones",
        "bandwidth": bandwidth,
        "batched_vs_serial_speedup": extras["ensemble"]["speedup"] if extras else None,
        "adjoint_over_forward": extras["adjoint"]["ratio"] if extras else None,
        "extras": extras,
    }
    print(json.dumps(out))


if __name__ == "__main__":
    main()
