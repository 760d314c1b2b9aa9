"""Autotuner of the tiled kernels' launch shape.

The tiled Verlet list and the force kernels have two free parameters: the tile height K
(cells per tile along z, i.e. the atoms each workgroup owns and the size of the neighbourhood
it stages in LDS) and the workgroup size.  Neither changes results (neighbour order within a
row follows the stencil, not the tile), only speed, and the best pair depends on density,
cell count and occupancy, so it is measured: each candidate is applied, the list rebuilt
and the force kernel timed with HIP events on the library stream.  The score is the force
time plus the build time spread over the expected steps between rebuilds.
"""
import itertools

DEFAULT_TILE_CELLS = (3, 4, 5, 6, 8)
DEFAULT_BLOCKS = (128, 256)


def _stat(ctx, name):
    n, t = ctx.kernel_stats(name)
    return t / n if n else 0.0


def tune_lj_system(system, tile_cells=DEFAULT_TILE_CELLS, blocks=DEFAULT_BLOCKS, force_reps=6,
                   steps_per_rebuild=6.0, verbose=False):
    """pick (tile cells, workgroup size) for an LJSystem-like object with .ctx, .vl, .atoms,
    .lj and .rebuild(); leaves the best setting applied and the list rebuilt.  Returns
    (best, table) with table = [(tile_cells, block, build_ms, force_ms, score), ...]."""
    ctx = system.ctx
    table = []
    was = ctx.profiling
    for K, block in itertools.product(tile_cells, blocks):
        system.vl.set_tiling(K, block)
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
        table.append((K, block, build, force, score))
        if verbose:
            print(f"tile_cells={K} block={block} staged={staged} build={build:.3f} ms force={force:.3f} ms",
                  flush=True)
    best = min(table, key=lambda r: r[4])
    system.vl.set_tiling(best[0], best[1])
    system.rebuild()
    ctx.set_profiling(was)
    return best, table
