"""Project autotuner for the tile kernel's launch shape.

Sweeps the knobs of the tile kernel on a representative batch and keeps the
fastest, timing each candidate with HIP events around the tile kernel
(profiling mode):

* the tile shape -- 256 leaves with one per lane (256-thread workgroups),
  256 leaves with two per lane (128 threads, two descents in flight per
  lane) or 128 leaves with one per lane (128 threads, twice the workgroups):
  the latency / occupancy trade of the greedy descent
  (kdh_set_leaves_per_lane, kdh_set_tile_leaves);
* the number of top-of-tree nodes staged in LDS per workgroup
  (0/256/512/1024, plus 2048 for rigid ICP);
* the workgroup count: one per tile, or a persistent grid of k workgroups
  per CU walking contiguous tile chunks;
* then, at the best shape, whether kdh_batch_align replays on one stream or
  on two (the pairs split in halves whose finish stages run beside the other
  half's tile launches), timed by whole batches (kdh_set_batch_streams).

The leaf size is a property of the trees, chosen when they are built
(TreeGenerator3D leaf_max_points).
"""
import numpy as np

N_CU = 256
LDS_CHOICES = (0, 256, 512, 1024, 2048)
SHAPES = ((1, 256), (2, 256), (1, 128))  # (leaves per lane, leaves per tile)


def tune(ctx, pairs, batch_align, ct, repeats=2, grids=(0, N_CU * 2, N_CU * 4, N_CU * 8), lds=LDS_CHOICES,
         shapes=SHAPES):
    """Returns (best config dict, all results).  Leaves the best config set.
    Each candidate is timed by one tile launch (all tiles of the batch) plus
    one finish launch: 128-leaf tiles cover the same leaves in twice the
    tiles, and the finish stage then sums twice the partial rows."""
    results = []
    ctx.set_profiling(True)
    try:
        for lpl, tile in shapes:
            for nodes in lds:
                if ct and nodes > 1024:
                    continue
                for grid in grids:
                    ctx.set_launch_config(grid=grid, lds_nodes=nodes, leaves_per_lane=lpl, tile_leaves=tile)
                    t, f = [], []
                    for _ in range(repeats):
                        batch_align(ctx, pairs)
                        tu, fu = ctx.kernel_breakdown()
                        t.append(tu)
                        f.append(fu)
                    results.append(dict(leaves_per_lane=lpl, tile_leaves=tile, lds_nodes=nodes, grid=grid,
                                        tile_us=float(np.median(t)), finish_us=float(np.median(f))))
    finally:
        ctx.set_profiling(False)
    # the tile shape also sets the finish stage's work (partial rows per pair)
    best = dict(min(results, key=lambda r: r["tile_us"] + r["finish_us"]))
    ctx.set_launch_config(grid=best["grid"], lds_nodes=best["lds_nodes"], leaves_per_lane=best["leaves_per_lane"],
                          tile_leaves=best["tile_leaves"])
    per_iter = {}
    for streams in (1, 2):
        ctx.set_launch_config(grid=best["grid"], lds_nodes=best["lds_nodes"], batch_streams=streams)
        t = []
        for _ in range(repeats + 2):
            batch_align(ctx, pairs)
            t.append(ctx.last_kernel_time()[0])
        per_iter[streams] = float(np.median(t[2:]))
    best["batch_streams"] = min(per_iter, key=per_iter.get)
    best["iteration_us"] = per_iter[best["batch_streams"]]
    ctx.set_launch_config(grid=best["grid"], lds_nodes=best["lds_nodes"], batch_streams=best["batch_streams"])
    return best, results
