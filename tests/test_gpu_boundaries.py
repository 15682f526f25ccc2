"""Size boundaries of the batched chain against the oracle, bit for bit:
render pairs of one triangle under, at and over a tile, and pairs whose
[bin][tile] histogram is exactly the largest one scanned by a single block
(128 tiles of 512 -> 65,536 entries) or one tile more (the three-kernel
scan).  Both execution forms."""
import math

import numpy as np
import pytest

from tetra3d_amd import _abi as A
from tetra3d_amd import workloads as W
from tetra3d_amd.scene import Camera, DirectionalLight, Mesh, Model, Scene, World
from tests.parity import compare_frames

f32 = np.float32
_GRID = {}


def _grid(n):
    if n not in _GRID:
        _GRID[n] = W.grid_mesh(W.SplitMix64(0x5eed + n), n, n, 1.0, 3.0, name="g%d" % n)
    return _GRID[n]


def _scene(triangles):
    """One grid model cut to exactly `triangles` triangles, seen from above by
    an orthographic camera that frames the whole grid."""
    n = int(math.ceil(math.sqrt(triangles / 2.0))) + 1
    g = _grid(n)
    mesh = Mesh("cut%d" % triangles, g.positions, g.normals, g.indices[:triangles], colors=g.colors, uvs=g.uvs)
    model = Model("grid", mesh)
    cam = Camera("cam", 1280, 1280)
    cam.perspective = False
    cam.ortho_scale = f32(n + 2)
    cam.far = f32(200)
    cam.set_local_position((n - 1) / 2.0, 50.0, (n - 1) / 2.0)
    cam.set_local_rotation(W.axis_angle_matrix((1, 0, 0), -math.pi / 2))
    sun = DirectionalLight("sun", 1, 1, 1, 1.0)
    sun.set_local_rotation(W.axis_angle_matrix((1, 0, 0), -1.0))
    world = World()
    world.ambient_light.energy = f32(0.3)
    return W.Workload("boundary_%d" % triangles, cam, Scene("boundary", world), [sun], [model])


CASES = [(512, 511), (512, 512), (512, 513), (2048, 2047), (2048, 2048), (2048, 2049),
         (512, 65536), (512, 65538)]


@pytest.mark.gpu
@pytest.mark.parametrize("mode", [A.EXEC_BATCHED, A.EXEC_PLAIN], ids=["batched", "plain"])
@pytest.mark.parametrize("case", CASES, ids=["tile%d_tris%d" % c for c in CASES])
def test_size_boundaries(gpu_ctx, oracle, case, mode):
    tile, tris = case
    wl = _scene(tris)
    assert wl.triangle_count == tris
    print("boundary", case, "mode", mode, flush=True)
    saved = gpu_ctx.tuning()
    gpu_ctx.set_tuning(triangle_tile=tile, assemble_split=1)
    try:
        g = gpu_ctx.render_workload(wl, exec_mode=mode)
    finally:
        gpu_ctx.L.t3d_set_tuning(gpu_ctx.ctx, saved)
    s = compare_frames(g, oracle.render_workload(wl))
    assert s["survivors"] > tris // 4


def _empty_cases():
    """(name, workload): no models; every model outside the frustum; a pair
    whose every triangle is back-facing next to a visible one."""
    base = _scene(512)
    yield "no_models", W.Workload("no_models", base.camera, base.scene, base.lights, [])
    away = _scene(512)
    away.camera.set_local_rotation(W.axis_angle_matrix((1, 0, 0), math.pi / 2))   # looking up, away from the grid
    yield "all_frustum_culled", away
    vis = _scene(512)
    flipped = _grid(18)
    back = Mesh("back", flipped.positions, flipped.normals, flipped.indices[:, ::-1].copy(), colors=flipped.colors)
    bm = Model("backfacing", back)
    bm.set_local_position(0, 0.5, 0)
    yield "backfacing_pair", W.Workload("backfacing_pair", vis.camera, vis.scene, vis.lights, vis.models + [bm])


@pytest.mark.gpu
@pytest.mark.parametrize("mode", [A.EXEC_BATCHED, A.EXEC_PLAIN], ids=["batched", "plain"])
def test_empty_and_fully_culled(gpu_ctx, oracle, mode):
    for name, wl in _empty_cases():
        print("empty case", name, "mode", mode, flush=True)
        g = gpu_ctx.render_workload(wl, exec_mode=mode)
        o = oracle.render_workload(wl)
        compare_frames(g, o)
        if name == "backfacing_pair":
            assert [p[3] for p in o.pairs][-1] == 0 and o.pairs[0][3] > 0
        else:
            assert g.survivors == 0
