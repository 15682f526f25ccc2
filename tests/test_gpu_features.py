"""Bit-exact GPU-vs-oracle parity on seeded scenes that combine the path's
options (see tests/feature_scenes.py), in both execution forms."""
import numpy as np
import pytest

from tetra3d_amd import _abi as A
from tests.feature_scenes import cell_lights_scene, feature_scene, many_bones_scene, pair_order_scene
from tests.parity import compare_frames

SEEDS = list(range(40))


@pytest.mark.gpu
@pytest.mark.parametrize("mode", [A.EXEC_BATCHED, A.EXEC_PLAIN], ids=["batched", "plain"])
def test_feature_scenes(gpu_ctx, oracle, mode):
    drawn = 0
    for seed in SEEDS:
        wl = feature_scene(seed)
        print("feature scene", seed, "mode", mode, flush=True)
        g = gpu_ctx.render_workload(wl, exec_mode=mode)
        s = compare_frames(g, oracle.render_workload(wl))
        drawn += s["survivors"]
    assert drawn > 0


@pytest.mark.gpu
@pytest.mark.parametrize("mode", [A.EXEC_BATCHED, A.EXEC_PLAIN], ids=["batched", "plain"])
def test_pair_order_ties(gpu_ctx, oracle, mode):
    """Render pairs at equal camera distances keep the reference's stable
    order (camera.go:1458-1467, 2050-2055)."""
    for seed in range(6):
        wl = pair_order_scene(seed)
        print("pair-order scene", seed, "mode", mode, flush=True)
        compare_frames(gpu_ctx.render_workload(wl, exec_mode=mode), oracle.render_workload(wl))


@pytest.mark.gpu
@pytest.mark.parametrize("mode", [A.EXEC_BATCHED, A.EXEC_PLAIN], ids=["batched", "plain"])
def test_many_bone_palettes(gpu_ctx, oracle, mode):
    """Palettes of 63-96 bones: staged in LDS up to 64 bones, read from
    device memory beyond (model.go:559-596 either way)."""
    for seed in range(2):
        wl = many_bones_scene(seed)
        print("many-bones scene", seed, "mode", mode, flush=True)
        s = compare_frames(gpu_ctx.render_workload(wl, exec_mode=mode), oracle.render_workload(wl))
        assert s["survivors"] > 0


def _same_nan_aware(g, o):
    """Pairs, draw order and indices equal; vertex fields bit for bit, every
    NaN equal to every NaN (a NaN's sign and payload follow the hardware)."""
    assert [p[:4] for p in g.pairs] == [p[:4] for p in o.pairs]
    np.testing.assert_array_equal(g.triangle_ids, o.triangle_ids)
    np.testing.assert_array_equal(g.indices, o.indices)
    for name in ("color", "depth"):
        x, y = getattr(g, name), getattr(o, name)
        both_nan = np.isnan(x) & np.isnan(y)
        bad = (x.view(np.uint32) != y.view(np.uint32)) & ~both_nan
        assert not bad.any(), (name, np.argwhere(bad)[:5])


@pytest.mark.gpu
@pytest.mark.parametrize("mode", [A.EXEC_BATCHED, A.EXEC_PLAIN], ids=["batched", "plain"])
def test_unculled_scenes(gpu_ctx, oracle, mode):
    """ProcessVertices with processOnlyVisible false for every pair (the
    BakeLighting and debug-view processing, model.go:600-960): the feature
    scenes, the closed-form scenes of the oracle pins, and a NaN key in a
    single-tile pair and in a many-tile pair."""
    from tests import test_oracle_pins as PINS
    from tetra3d_amd.scene import Material, Model, Scene
    from tetra3d_amd import workloads as W
    scenes = []
    for seed in range(0, 40, 3):
        wl = feature_scene(seed)
        wl.camera.process_only_visible = False
        scenes.append((wl, False))
    for nan in (False, True):
        m = PINS._unculled_scene()
        if nan:
            pos = m.mesh.positions.copy()
            pos[12, 2] = np.nan
            nm = PINS.tri_mesh(pos)
            nm.mesh_parts[0].material = m.mesh.mesh_parts[0].material
            m = PINS.model(nm, backface=True)
        cam = PINS.ortho_camera()
        cam.process_only_visible = False
        scenes.append((W.Workload("unculled_pins_%d" % nan, cam, Scene("s", PINS.world()), [], [m]), nan))
    big = W.icosphere(5, 1.5, name="nan_big")        # 20,480 triangles: ten tiles of 2,048
    pos = big.positions.copy()
    pos[np.random.default_rng(9).integers(0, len(pos), 40)] = np.nan
    bm = W.icosphere(5, 1.5, name="nan_big2")
    bm.positions = pos
    bm.mesh_parts[0].material = Material("nan")
    mb = Model("nan_big", bm)
    mb.frustum_culling = False
    mb.set_local_position(0, 0, -4)
    cam = PINS.persp_camera(far=20.0)
    cam.process_only_visible = False
    scenes.append((W.Workload("unculled_nan_tiles", cam, Scene("s", PINS.world(0.5)), [], [mb]), True))
    drawn = 0
    for wl, nan in scenes:
        print("unculled", wl.name, "mode", mode, flush=True)
        g, o = gpu_ctx.render_workload(wl, exec_mode=mode), oracle.render_workload(wl)
        if nan:
            _same_nan_aware(g, o)
        else:
            compare_frames(g, o)
        drawn += o.survivors
    assert drawn > 0


@pytest.mark.gpu
@pytest.mark.parametrize("mode", [A.EXEC_BATCHED, A.EXEC_PLAIN], ids=["batched", "plain"])
@pytest.mark.parametrize("n", [64, 70, 200])
def test_many_lights(gpu_ctx, oracle, mode, n):
    """A render pair may see any number of lights: up to 64 are staged in
    LDS by the vertex pass, more are read from device memory in list order
    (camera.go:1610-1644); point lights around the models, a directional
    light and a LightGroup of as many lights on one model."""
    from tetra3d_amd import workloads as W
    rng = np.random.default_rng(n)
    wl = feature_scene(3)
    wl.camera.process_only_visible = True
    lights = [W._directional("sun", (0.3, -1.0, 0.2))]
    lights += [W._point("p%d" % i, rng.uniform(-6, 6, 3), float(rng.uniform(4, 30))) for i in range(n - 1)]
    wl.lights = lights
    wl.models[0].light_group = lights[::-1]
    print("lights", n, "mode", mode, flush=True)
    s = compare_frames(gpu_ctx.render_workload(wl, exec_mode=mode), oracle.render_workload(wl))
    assert s["survivors"] > 0


@pytest.mark.gpu
@pytest.mark.parametrize("mode", [A.EXEC_BATCHED, A.EXEC_PLAIN], ids=["batched", "plain"])
@pytest.mark.parametrize("many", [False, True], ids=["few", "many"])
def test_cell_lights(gpu_ctx, oracle, mode, many):
    """A LightVolume whose cells name lights (light.go:646-660): each of a
    cell's lights lights the vertex before the volume's colour, unskipped and
    whatever its visibility, with beginModel for the model at hand; rigid and
    skinned models in every shading mode, the volume in the scene lights and
    in LightGroups; with 80 cell lights a pair's list passes the 64 lights
    the vertex pass stages in LDS."""
    for seed in range(4):
        wl = cell_lights_scene(seed, many)
        print("cell-lights scene", seed, "many", many, "mode", mode, flush=True)
        s = compare_frames(gpu_ctx.render_workload(wl, exec_mode=mode), oracle.render_workload(wl))
        assert s["survivors"] > 0


@pytest.mark.gpu
def test_cell_lights_follow_list_changes(gpu_ctx, oracle):
    """Cell lists changed between renders (SetUseLights, ClearUseLights)
    reach the device before the next render."""
    wl = cell_lights_scene(7)
    vol = wl.lights[1]
    for step in range(3):
        if step == 1:
            for c in range(0, len(vol.cells), 2):
                vol.clear_use_lights(c)
        elif step == 2:
            vol.set_use_lights(0, *vol.use_lights_render_set[:2])
        print("cell-lights change", step, flush=True)
        compare_frames(gpu_ctx.render_workload(wl), oracle.render_workload(wl))
