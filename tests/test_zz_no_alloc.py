"""t3d_render_noalloc (the render that allocates no device memory) against
t3d_render through the C-ABI: two calls in a row into the same caller-owned
device arrays, the second with a smaller frame, each bit-identical to the
allocating entry point's frame; and its two refusals."""
import ctypes as C

import numpy as np
import pytest

from tetra3d_amd import _abi as A
from tetra3d_amd import runtime
from tetra3d_amd import workloads as W
from tetra3d_amd.runtime import Context, T3DError

HIP_MEMCPY_D2H = 2


def _hip():
    runtime.lib()                       # loads libamdhip64 as the library's dependency
    L = C.CDLL("libamdhip64.so.7")      # the runtime libt3d.so already uses (same soname)
    L.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
    L.hipFree.argtypes = [C.c_void_p]
    L.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
    return L


class DeviceArrays:
    """Caller-owned output arrays: colour and depth vertices and indices."""

    def __init__(self, hip, capacity):
        self.hip, self.capacity = hip, capacity
        self.ptrs = []
        for nbytes in (48 * capacity, 48 * capacity, 4 * capacity):
            p = C.c_void_p()
            assert hip.hipMalloc(C.byref(p), nbytes) == 0
            self.ptrs.append(p.value)

    def read(self):
        c = np.empty((self.capacity, 12), dtype=np.float32)
        d = np.empty((self.capacity, 12), dtype=np.float32)
        i = np.empty(self.capacity, dtype=np.uint32)
        for arr, p in zip((c, d, i), self.ptrs):
            assert self.hip.hipMemcpy(arr.ctypes.data, p, arr.nbytes, HIP_MEMCPY_D2H) == 0
        return c, d, i

    def free(self):
        for p in self.ptrs:
            self.hip.hipFree(p)
        self.ptrs = []


def _max_vertices(wl):
    return 3 * sum(m.mesh.triangle_count for m in wl.models if m.mesh is not None)


def _noalloc_frame(ctx, wl, arrays, frame, mode):
    inp = ctx.inputs(wl.camera, wl.scene, wl.lights, wl.models, mode)
    ctx.render_noalloc(inp, *arrays.ptrs, arrays.capacity, frame=frame)
    pairs = ctx.pairs(frame)
    c, d, i = arrays.read()
    sel = np.concatenate([np.arange(p[4], p[4] + 3 * p[3]) for p in pairs] or [np.zeros(0, np.int64)])
    return pairs, c[sel], d[sel], i[sel]


@pytest.mark.gpu
@pytest.mark.parametrize("mode", [A.EXEC_BATCHED, A.EXEC_PLAIN], ids=["batched", "plain"])
def test_noalloc_matches_render_twice_shrinking(mode):
    big = W.mixed_scene(count=12)
    small = W.mixed_scene(seed=11, count=4)
    hip = _hip()
    with Context(0) as ctx:
        ctx.poison_outputs(0x7FC0DEAD)
        frame = ctx.new_frame()
        arrays = DeviceArrays(hip, _max_vertices(big))
        try:
            for name, wl in (("mixed_12", big), ("mixed_4", small)):
                print("noalloc", name, "mode", mode, "tris", wl.triangle_count, flush=True)
                ref = ctx.render_workload(wl, exec_mode=mode)
                pairs, c, d, i = _noalloc_frame(ctx, wl, arrays, frame, mode)
                assert [p[:4] for p in pairs] == [p[:4] for p in ref.pairs]
                assert ref.survivors > 0
                np.testing.assert_array_equal(c.view(np.uint32), ref.color.view(np.uint32))
                np.testing.assert_array_equal(d.view(np.uint32), ref.depth.view(np.uint32))
                np.testing.assert_array_equal(i, ref.indices)
                print("noalloc", name, "bit-identical,", ref.survivors, "drawn triangles", flush=True)
        finally:
            arrays.free()


@pytest.mark.gpu
def test_noalloc_cell_lights():
    """A LightVolume whose cells name 80 lights (past the 64 the vertex pass
    stages in LDS), skinned and rigid models: the no-alloc render matches
    t3d_render bit for bit."""
    from tests.feature_scenes import cell_lights_scene
    wl = cell_lights_scene(2, many=True)
    hip = _hip()
    with Context(0) as ctx:
        frame = ctx.new_frame()
        arrays = DeviceArrays(hip, _max_vertices(wl))
        try:
            print("noalloc cell lights", flush=True)
            ref = ctx.render_workload(wl)
            pairs, c, d, i = _noalloc_frame(ctx, wl, arrays, frame, A.EXEC_BATCHED)
            assert [p[:4] for p in pairs] == [p[:4] for p in ref.pairs]
            assert ref.survivors > 0
            np.testing.assert_array_equal(c.view(np.uint32), ref.color.view(np.uint32))
            np.testing.assert_array_equal(d.view(np.uint32), ref.depth.view(np.uint32))
            np.testing.assert_array_equal(i, ref.indices)
        finally:
            arrays.free()


@pytest.mark.gpu
def test_noalloc_refusals():
    hip = _hip()
    wl = W.mixed_scene(count=6)
    big = W.heightfield_scene(n=1024)       # ~125 MB of tables and scratch: past the static array
    with Context(0) as ctx:
        ctx.prepare(wl)
        ctx.prepare(big)
        frame = ctx.new_frame()
        arrays = DeviceArrays(hip, 3)
        try:
            inp = ctx.inputs(wl.camera, wl.scene, wl.lights, wl.models)
            with pytest.raises(T3DError) as e:
                ctx.render_noalloc(inp, *arrays.ptrs, arrays.capacity, frame=frame)
            assert e.value.status == A.T3D_ERR_INVALID_ARGUMENT and "outputs hold" in str(e.value)
            inp = ctx.inputs(big.camera, big.scene, big.lights, big.models)
            with pytest.raises(T3DError) as e:
                ctx.render_noalloc(inp, *arrays.ptrs, arrays.capacity, frame=frame)
            assert e.value.status == A.T3D_ERR_BUDGET and "static scratch" in str(e.value)
            # the context renders exactly afterwards
            ref = ctx.render_workload(wl)
            assert ref.survivors > 0
        finally:
            arrays.free()
