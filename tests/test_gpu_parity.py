"""GPU parity: the HIP path through the C-ABI against the CPU oracle on the
same seeded scenes (sizes the oracle finishes in seconds)."""
import numpy as np
import pytest

from tetra3d_amd import _abi as A
from tetra3d_amd import workloads as W
from tests.parity import compare_frames

SCENES = {
    "icosphere": lambda: W.icosphere_scene(),
    "cubes": lambda: W.cubes_scene(count=1024),
    "heightfield_128": lambda: W.heightfield_scene(n=128),
    "mixed_48": lambda: W.mixed_scene(count=48),
}


@pytest.mark.gpu
@pytest.mark.parametrize("variant", [(0, 0), (1, 0), (2, 0), (2, 1), (2, 2), (2, 3)],
                         ids=["fused_staged", "fused_direct", "split_staged", "split_direct", "split_staged_nt",
                              "split_direct_nt"])
@pytest.mark.parametrize("mode", [A.EXEC_BATCHED, A.EXEC_PLAIN], ids=["batched", "plain"])
@pytest.mark.parametrize("scene", list(SCENES))
def test_scene_parity(gpu_ctx, oracle, scene, mode, variant):
    wl = SCENES[scene]()
    print("parity", scene, "mode", mode, "variant", variant, "tris", wl.triangle_count, flush=True)
    gpu_ctx.set_tuning(emit_variant=variant[0], store_mode=variant[1])
    try:
        g = gpu_ctx.render_workload(wl, exec_mode=mode)
    finally:
        gpu_ctx.set_tuning(emit_variant=2, store_mode=2)
    o = oracle.render_workload(wl)
    s = compare_frames(g, o)
    print(s, flush=True)
    assert s["survivors"] > 0


@pytest.mark.gpu
def test_device_float_primitives_match_host(gpu_ctx):
    """Division, sqrt and round on the device round exactly as IEEE float32 on the host."""
    rng = np.random.default_rng(7)
    x = (rng.standard_normal(1 << 16) * 10.0 ** rng.integers(-20, 20, 1 << 16)).astype(np.float32)
    y = (rng.standard_normal(1 << 16) * 10.0 ** rng.integers(-20, 20, 1 << 16)).astype(np.float32)
    out = gpu_ctx.selftest_fp(x, y)
    with np.errstate(all="ignore"):
        np.testing.assert_array_equal(out[0].view(np.uint32), (x / y).view(np.uint32))
        np.testing.assert_array_equal(out[1].view(np.uint32), np.sqrt(np.abs(x)).view(np.uint32))
        np.testing.assert_array_equal(out[2].view(np.uint32), (x * y + x).view(np.uint32))
        # round half away from zero, checked where float64 arithmetic is exact
        ok = np.abs(x) < 2 ** 22
        expect = (np.sign(x[ok].astype(np.float64)) * np.floor(np.abs(x[ok].astype(np.float64)) + 0.5)).astype(np.float32)
        np.testing.assert_array_equal(out[3][ok], expect)
