"""Project autotuner for the tile kernel's launch shape.

Sweeps the knobs of kdh_set_launch_config -- the number of top-of-tree nodes
staged in LDS (0/256/512/1024, plus 2048 for rigid ICP) and the workgroup
count (one per tile, or a persistent grid of k workgroups per CU) -- on a
representative batch, timing each candidate with HIP events around the tile
kernel (profiling mode), and keeps the fastest.  The tile size is fixed at
256 moving leaves (one workgroup pass); the leaf size is a property of the
trees, chosen when they are built.
"""
import numpy as np

N_CU = 256
LDS_CHOICES = (0, 256, 512, 1024, 2048)


def tune(ctx, pairs, batch_align, ct, repeats=2, grids=(0, N_CU * 2, N_CU * 4, N_CU * 8)):
    """Returns (best config dict, all results).  Leaves the best config set."""
    results = []
    ctx.set_profiling(True)
    try:
        for lds in LDS_CHOICES:
            if ct and lds > 1024:
                continue
            for grid in grids:
                ctx.set_launch_config(grid=grid, lds_nodes=lds)
                t = []
                for _ in range(repeats):
                    batch_align(ctx, pairs)
                    t.append(ctx.kernel_breakdown()[0])
                results.append(dict(lds_nodes=lds, grid=grid, tile_us=float(np.median(t))))
    finally:
        ctx.set_profiling(False)
    best = min(results, key=lambda r: r["tile_us"])
    ctx.set_launch_config(grid=best["grid"], lds_nodes=best["lds_nodes"])
    return best, results
