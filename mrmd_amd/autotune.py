"""Autotuners of the tiled kernels' launch shape (single systems and ensembles).

The tiled Verlet list and the force kernels have three free parameters: the tile shape
(a x a cell columns by K cells along z, i.e. the atoms each workgroup owns and the size of
the neighbourhood it stages in LDS) and the workgroup size.  Neither changes results (neighbour order within a
row follows the stencil, not the tile), only speed, and the best pair depends on density,
cell count and occupancy, so it is measured: each candidate is applied, the list rebuilt
and the force kernel timed with HIP events on the library stream.  The score is the force
time plus the build time spread over the expected steps between rebuilds.
"""
import itertools

DEFAULT_SHAPES = ((1, 6), (2, 2), (2, 3), (2, 4), (3, 1), (3, 2))  # (columns, cells)
DEFAULT_BLOCKS = (192, 256)


def _stat(ctx, name):
    n, t = ctx.kernel_stats(name)
    return t / n if n else 0.0


def tune_lj_system(system, shapes=DEFAULT_SHAPES, blocks=DEFAULT_BLOCKS, force_reps=6, steps_per_rebuild=6.0,
                   verbose=False):
    """pick (tile columns, tile cells, workgroup size) for an LJSystem-like object with .ctx,
    .vl, .atoms, .lj and .rebuild(); leaves the best setting applied and the list rebuilt.
    Returns (best, table) with table = [(columns, cells, block, build_ms, force_ms, score)]."""
    ctx = system.ctx
    table = []
    was = ctx.profiling
    for (a, K), block in itertools.product(shapes, blocks):
        system.vl.set_tiling(K, block, a)
        ctx.set_profiling(True)
        ctx.reset_stats()
        system.rebuild()
        build = _stat(ctx, "verlet_build")
        for _ in range(force_reps):
            system.atoms.set_force(0.0)
            system.lj.apply(system.atoms, system.vl)
        ctx.synchronize()
        force = _stat(ctx, "lj_force")
        ctx.set_profiling(False)
        staged = system.vl.tiling()["max_staged"]
        score = force + build / steps_per_rebuild
        table.append((a, K, block, build, force, score))
        if verbose:
            print(f"columns={a} tile_cells={K} block={block} staged={staged} build={build:.3f} ms "
                  f"force={force:.3f} ms", flush=True)
    best = min(table, key=lambda r: r[5])
    system.vl.set_tiling(best[1], best[2], best[0])
    system.rebuild()
    ctx.set_profiling(was)
    return best, table


ENSEMBLE_TARGETS = (1000, 1400, 1800)  # staged records per workgroup a replica's tile height aims at
ENSEMBLE_BLOCKS = (128, 256)


def tune_ensemble(ensemble, targets=ENSEMBLE_TARGETS, blocks=ENSEMBLE_BLOCKS, steps=20, graph_steps=10,
                  verbose=False):
    """pick the ensemble's launch shape: the target number of staged records per workgroup
    (which sets every replica's tile height K from its density, i.e. the atoms per workgroup
    and the cell-tile shape) and the workgroup size.  Each candidate runs `steps` steps of the
    live ensemble (after a short untimed run that rebuilds its lists and graphs) and the
    fastest wall time wins; results do not depend on the shape, so tuning advances the
    ensemble's state like any other steps.  Leaves the best setting applied and returns
    ({"target_staged", "block"}, table of (target, block, seconds))."""
    import time
    ctx = ensemble.ctx
    table = []
    for target in targets:
        for block in blocks:
            ensemble.set_tiling(target, block)
            ensemble.run(graph_steps, graph_steps)
            ctx.synchronize()
            t0 = time.perf_counter()
            ensemble.run(steps, graph_steps)
            ctx.synchronize()
            dt = time.perf_counter() - t0
            table.append((target, block, dt))
            if verbose:
                print(f"target_staged={target} block={block} {dt * 1e3 / steps:.3f} ms/step", flush=True)
    best = min(table, key=lambda r: r[2])
    ensemble.set_tiling(best[0], best[1])
    return {"target_staged": best[0], "block": best[1]}, table
