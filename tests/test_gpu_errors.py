"""Invalid inputs fail with a status and a message, allocate nothing that
leaks and launch nothing, and the context renders correctly afterwards (the
reference panics or errors where this library returns a status:
camera.go:389-391, mesh.go:681, 2279).  Every case is refused by host-side
validation, before any kernel."""
import ctypes as C

import numpy as np
import pytest

from tetra3d_amd import _abi as A
from tetra3d_amd import workloads as W
from tetra3d_amd.runtime import T3DError
from tests.feature_scenes import many_bones_scene
from tests.parity import compare_frames


def _refused(ctx, wl, mutate, status):
    inp = ctx.inputs(wl.camera, wl.scene, wl.lights, wl.models)
    mutate(inp.render)
    with pytest.raises(T3DError) as e:
        ctx.render_inputs(inp)
    assert e.value.status == status, (e.value.status, status)


@pytest.mark.gpu
def test_invalid_render_inputs_are_refused(gpu_ctx, oracle):
    wl = W.icosphere_scene()
    gpu_ctx.prepare(wl)
    print("invalid render inputs", flush=True)

    def custom_depth(r):
        r.models[0].materials[0].present = 1
        r.models[0].materials[0].visible = 1
        r.models[0].materials[0].custom_depth_function = 1

    cases = [
        (lambda r: setattr(r.models[0], "mesh", 0x7FFFFFFF), A.T3D_ERR_NOT_FOUND),
        (lambda r: setattr(r.camera[0], "width", 0), A.T3D_ERR_INVALID_ARGUMENT),
        (lambda r: setattr(r.models[0], "has_vertex_transform_function", 1), A.T3D_ERR_UNSUPPORTED),
        (lambda r: setattr(r.models[0], "has_vertex_clip_function", 1), A.T3D_ERR_UNSUPPORTED),
        (lambda r: setattr(r.lights[0], "type", 99), A.T3D_ERR_INVALID_ARGUMENT),
        (lambda r: setattr(r.models[0], "skinned", 1), A.T3D_ERR_INVALID_ARGUMENT),   # no skin data
        (custom_depth, A.T3D_ERR_UNSUPPORTED),
    ]
    for mutate, status in cases:
        _refused(gpu_ctx, wl, mutate, status)
    # a bone index past the model's palette
    sk = many_bones_scene(0)
    gpu_ctx.prepare(sk)
    _refused(gpu_ctx, sk, lambda r: setattr(r.models[3], "bone_count", 10), A.T3D_ERR_INVALID_ARGUMENT)

    # a skinned model with the unbillboarded depth mode (the reference reads a depth it never wrote)
    def unbb_skinned(r):
        m = next(i for i in range(r.model_count) if r.models[i].skinned)
        r.models[m].materials[0].present = 1
        r.models[m].materials[0].visible = 1
        r.models[m].materials[0].depth_mode = A.DEPTH_MODE_UNBILLBOARDED
    _refused(gpu_ctx, sk, unbb_skinned, A.T3D_ERR_UNSUPPORTED)

    # dynamic batches naming no model, the batched model itself, or a part the owner does not have
    def batched(owner, part):
        def f(r):
            r.models[1].dynamic_batched = 1
            r.models[1].batch_owner = owner
            r.models[1].batch_part = part
        return f
    for owner, part in (len(sk.models) + 5, 0), (1, 0), (0, 99):
        _refused(gpu_ctx, sk, batched(owner, part), A.T3D_ERR_INVALID_ARGUMENT)
    # the context still renders both scenes exactly
    compare_frames(gpu_ctx.render_workload(wl), oracle.render_workload(wl))
    compare_frames(gpu_ctx.render_workload(sk), oracle.render_workload(sk))


@pytest.mark.gpu
def test_invalid_cell_lights_are_refused(gpu_ctx, oracle):
    """Cell light lists that do not fit the volume, a volume light with fewer
    use_lights than its cells name, and a LightVolume among a volume's cell
    lights are refused; the scene renders afterwards."""
    from tests.feature_scenes import cell_lights_scene
    from tetra3d_amd.scene import LightVolume
    wl = cell_lights_scene(1)
    gpu_ctx.prepare(wl)
    vol = wl.lights[1]
    print("invalid cell lights", flush=True)
    U = C.POINTER(C.c_uint32)
    n = len(vol.cells)
    bad_first = np.zeros(n + 1, dtype=np.uint32)
    bad_first[0] = 1
    idx = np.zeros(4, dtype=np.uint32)
    assert gpu_ctx.L.t3d_volume_set_cell_lights(gpu_ctx.ctx, vol.handle, bad_first.ctypes.data_as(U),
                                                idx.ctypes.data_as(U)) == A.T3D_ERR_INVALID_ARGUMENT
    dec = np.zeros(n + 1, dtype=np.uint32)
    dec[1] = 3
    assert gpu_ctx.L.t3d_volume_set_cell_lights(gpu_ctx.ctx, vol.handle, dec.ctypes.data_as(U),
                                                idx.ctypes.data_as(U)) == A.T3D_ERR_INVALID_ARGUMENT
    assert gpu_ctx.L.t3d_volume_set_cell_lights(gpu_ctx.ctx, 0x7FFFFFFF, None, None) == A.T3D_ERR_NOT_FOUND

    def fewer(r):
        r.lights[1].use_light_count = 1
    _refused(gpu_ctx, wl, fewer, A.T3D_ERR_INVALID_ARGUMENT)

    inner = LightVolume("inner", (-1, -1, -1), (1, 1, 1), (1, 1, 1))
    gpu_ctx.upload_volume(inner)
    keep = []

    def nested_volume(r):
        arr = (A.LightDesc * r.lights[1].use_light_count)()
        for k in range(r.lights[1].use_light_count):
            arr[k] = r.lights[1].use_lights[k]
        arr[0] = inner.desc()
        keep.append(arr)
        r.lights[1].use_lights = arr
    _refused(gpu_ctx, wl, nested_volume, A.T3D_ERR_UNSUPPORTED)
    gpu_ctx.sync_cell_lights(vol)
    compare_frames(gpu_ctx.render_workload(wl), oracle.render_workload(wl))


@pytest.mark.gpu
def test_invalid_tuning_is_refused(gpu_ctx, oracle):
    print("invalid tuning", flush=True)
    before = gpu_ctx.tuning()
    for kw in [dict(triangle_tile=100), dict(triangle_tile=4096), dict(store_mode=5), dict(scatter_waves=2),
               dict(assemble_split=3), dict(vertex_block=320), dict(cull_block=128), dict(xcd_swizzle=2)]:
        with pytest.raises(T3DError) as e:
            gpu_ctx.set_tuning(**kw)
        assert e.value.status == A.T3D_ERR_INVALID_ARGUMENT, kw
    after = gpu_ctx.tuning()
    assert bytes(before) == bytes(after)          # a refused tuning changes nothing
    wl = W.icosphere_scene()
    compare_frames(gpu_ctx.render_workload(wl), oracle.render_workload(wl))
