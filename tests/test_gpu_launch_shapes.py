"""Launch shapes of the tile kernel give the same bits: one workgroup per
tile, or a persistent grid of fewer workgroups that each walk a contiguous
chunk of tiles (with the XCD mapping when the grid is a multiple of 8, and a
grid below 8 without it), at every staging depth.  Which workgroup runs a tile
changes neither its decisions nor its partial row, so whole batches agree
byte for byte; poses also agree with the oracle.  GPU."""
import numpy as np
import pytest

from kd_slam2_amd import Context, DeviceTree, TreeGenerator3D, batch_align, synth
from oracle import pyoracle as O

from .workloads import icp_params

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def shape_batch():
    ctx = Context(0)
    g = TreeGenerator3D(leaf_max_points=16)
    ws, trees = [], []
    # scans of 3k-40k points: 1 to 10 tiles per pair, 27 tiles in all
    for i, n in enumerate([3000, 40000, 9000, 17000, 5000, 26000]):
        w = synth.config1(seed=5200 + i, n_map=3 * n, n_scan=n, rot_deg=5.0, trans_m=1.0, uniform_init=True)
        w["fixed"], w["moving"] = g.makeTree(w["map"]), g.makeTree(w["scan"])
        ws.append(w)
        trees.append((DeviceTree(ctx, w["fixed"]), DeviceTree(ctx, w["moving"])))
    p = icp_params()
    pairs = [dict(fixed=F, moving=M, ct=0, X0=w["X0"], Xf=w["Xf"], params=p) for w, (F, M) in zip(ws, trees)]
    yield ctx, ws, pairs, p
    for F, M in trees:
        F.close()
        M.close()
    ctx.close()


def _bytes(res):
    return [(r["X"].tobytes(), r["result"], r["iterations"], repr(sorted(r["last"].items()))) for r in res]


@pytest.mark.parametrize("lds", [0, 256, 1024])
def test_persistent_grids_match_one_workgroup_per_tile(shape_batch, lds):
    ctx, ws, pairs, p = shape_batch
    n_tiles = sum((w["moving"].num_leaves + 255) // 256 for w in ws)
    assert n_tiles > 24  # the grids below are persistent for this batch
    ctx.set_launch_config(grid=0, lds_nodes=lds, leaves_per_lane=1, tile_leaves=256)
    ref = batch_align(ctx, pairs)
    for grid in (4, 8, 16, 24):
        ctx.set_launch_config(grid=grid, lds_nodes=lds, leaves_per_lane=1, tile_leaves=256)
        assert _bytes(batch_align(ctx, pairs)) == _bytes(ref), grid
    ctx.set_launch_config(grid=0, lds_nodes=256, leaves_per_lane=1, tile_leaves=256)
    for i, w in enumerate(ws):
        Xo, _, _, res_o = O.compute(w["fixed"], w["moving"], w["X0"], w["Xf"], O.from_kdh_params(p))
        dt, dr = synth.pose_error(ref[i]["X"], Xo)
        assert ref[i]["result"] == res_o and dt < 1e-3 and dr < 1e-3, (i, dt, dr)
