"""Bit-exact GPU-vs-oracle parity on seeded scenes that combine the path's
options (see tests/feature_scenes.py), in both execution forms."""
import pytest

from tetra3d_amd import _abi as A
from tests.feature_scenes import feature_scene, many_bones_scene, pair_order_scene
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
