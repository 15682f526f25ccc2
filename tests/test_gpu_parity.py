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
@pytest.mark.parametrize("store_mode", [0, 1, 2, 3, 4, 6, 8, 10, 12, 14, 16, 18, 20, 22, 34, 38, 50, 54, 80, 82, 114],
                         ids=["staged", "direct", "staged_nt", "direct_nt", "half", "half_nt", "pipe", "pipe_nt",
                              "pipe_half", "pipe_half_nt", "w4", "w4_nt", "w4_half", "w4_half_nt", "one_nt",
                              "one_half_nt", "w4_one_nt", "w4_one_half_nt", "w4_noswz", "w4_nt_noswz",
                              "w4_one_nt_noswz"])
@pytest.mark.parametrize("mode", [A.EXEC_BATCHED, A.EXEC_PLAIN], ids=["batched", "plain"])
@pytest.mark.parametrize("scene", list(SCENES))
def test_scene_parity(gpu_ctx, oracle, scene, mode, store_mode):
    wl = SCENES[scene]()
    print("parity", scene, "mode", mode, "store_mode", store_mode, "tris", wl.triangle_count, flush=True)
    gpu_ctx.set_tuning(store_mode=store_mode)
    try:
        g = gpu_ctx.render_workload(wl, exec_mode=mode)
    finally:
        gpu_ctx.set_tuning(store_mode=2)
    o = oracle.render_workload(wl)
    s = compare_frames(g, o)
    print(s, flush=True)
    assert s["survivors"] > 0


SHAPES = [  # (triangle_tile, scatter_waves, assemble_split, xcd_swizzle, vertex_split, cull_block)
    (512, 1, 1, 0, 0, 256), (512, 4, 8, 1, 1, 1024), (1024, 4, 1, 1, 0, 512), (1024, 4, 4, 0, 1, 256),
    (2048, 1, 2, 1, 1, 512), (2048, 4, 32, 1, 0, 1024), (64, 4, 1, 1, 1, 1024),
]


@pytest.mark.gpu
@pytest.mark.parametrize("shape", SHAPES, ids=["t%d_w%d_s%d_x%d_v%d_c%d" % s for s in SHAPES])
@pytest.mark.parametrize("scene", ["heightfield_128", "mixed_48"])
def test_launch_shapes_parity(gpu_ctx, oracle, scene, shape):
    """Every tile size, scatter width, assembly split, block order and vertex
    record layout gives the same frame."""
    wl = SCENES[scene]()
    tile, waves, split, swz, vsplit, cblock = shape
    print("shape parity", scene, shape, flush=True)
    saved = gpu_ctx.tuning()
    gpu_ctx.set_tuning(triangle_tile=tile, scatter_waves=waves, assemble_split=split, xcd_swizzle=swz,
                       vertex_split=vsplit, cull_block=cblock)
    try:
        g = gpu_ctx.render_workload(wl)
    finally:
        gpu_ctx.L.t3d_set_tuning(gpu_ctx.ctx, saved)
    s = compare_frames(g, oracle.render_workload(wl))
    assert s["survivors"] > 0


@pytest.mark.gpu
def test_async_frames_overlap_and_match(gpu_ctx, oracle):
    """Two frames rendered asynchronously back to back (the second's host
    work overlapping the first's kernels) give the synchronous results."""
    from tetra3d_amd.runtime import T3DError
    wa, wb = SCENES["mixed_48"](), SCENES["cubes"]()
    gpu_ctx.prepare(wa)
    gpu_ctx.prepare(wb)
    ia, ib = gpu_ctx.inputs(wa.camera, wa.scene, wa.lights, wa.models), gpu_ctx.inputs(wb.camera, wb.scene, wb.lights, wb.models)
    fa, fb = gpu_ctx.new_frame(), gpu_ctx.new_frame()
    for rep in range(3):
        print("async round", rep, flush=True)
        gpu_ctx.render_async(ia, fa)
        gpu_ctx.render_async(ib, fb)
        with pytest.raises(T3DError) as e:
            gpu_ctx.pairs(fb)                    # in flight until waited for
        assert e.value.status == A.T3D_ERR_NOT_READY
        gpu_ctx.wait(fa)
        gpu_ctx.wait(fb)
        compare_frames(gpu_ctx.download(frame=fa), oracle.render_workload(wa))
        compare_frames(gpu_ctx.download(frame=fb), oracle.render_workload(wb))


FULL = {
    "icosphere_642v": lambda: W.icosphere_scene(),
    "cubes_4096": lambda: W.cubes_scene(),
    "heightfield_1024": lambda: W.heightfield_scene(),
    "mixed_512": lambda: W.mixed_scene(),
    "crowd_256": lambda: W.crowd_scene(),
    "icosphere_ref_1272v": lambda: W.reference_icosphere_scene(),
}


@pytest.mark.gpu
@pytest.mark.parametrize("scene", list(FULL))
def test_full_size_parity(gpu_ctx, oracle, scene):
    """The measured configurations at their full sizes, autotuned launch
    shape included, bit-exact against the oracle (a few seconds each)."""
    wl = FULL[scene]()
    print("full-size parity", scene, "tris", wl.triangle_count, flush=True)
    saved = gpu_ctx.tuning()
    try:
        if scene == "mixed_512":
            gpu_ctx.prepare(wl)
            gpu_ctx.autotune(gpu_ctx.inputs(wl.camera, wl.scene, wl.lights, wl.models), reps=1)
        g = gpu_ctx.render_workload(wl)
    finally:
        gpu_ctx.L.t3d_set_tuning(gpu_ctx.ctx, saved)
    s = compare_frames(g, oracle.render_workload(wl))
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
