"""Autotuner of the tiled kernels' launch shape.

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
