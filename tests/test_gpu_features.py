"""Bit-exact GPU-vs-oracle parity on seeded scenes that combine the path's
options (see tests/feature_scenes.py), in both execution forms."""
import pytest

from tetra3d_amd import _abi as A
from tests.feature_scenes import feature_scene
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
