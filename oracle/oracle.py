"""ctypes binding of the CPU oracle (oracle/libt3d_oracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, by __graft_entry__.smoke() and
by bench.py's CPU-baseline leg, always as the checker.  The product package
never imports this module.
"""
import ctypes as C
import time
import os
import subprocess

import numpy as np

from tetra3d_amd import _abi as A
from tetra3d_amd.scene import RenderInputs, with_batched

HERE = os.path.dirname(os.path.abspath(__file__))
# T3D_ORACLE_LIB points the checker at another build of the same sources (a
# copy with a planted mistake, to show that a test fails on it)
LIB_PATH = os.environ.get("T3D_ORACLE_LIB") or os.path.join(HERE, "libt3d_oracle.so")


class OracleResult(C.Structure):
    _fields_ = [
        ("pair_count", C.c_uint32),
        ("pairs", C.POINTER(A.PairResult)),
        ("vertex_count", C.c_uint64),
        ("color", C.POINTER(A.Vertex)),
        ("depth", C.POINTER(A.Vertex)),
        ("indices", C.POINTER(C.c_uint32)),
        ("triangle_ids", C.POINTER(C.c_uint32)),
        ("depth_keys", C.POINTER(C.c_float)),
        ("normals", C.POINTER(A.Vertex)),
    ]


def build():
    subprocess.run(["make", "-s", "-C", HERE], check=True)


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        L = C.CDLL(LIB_PATH)
        M4 = C.POINTER(A.Mat4)
        FP = C.POINTER(C.c_float)
        L.t3do_create.restype = C.c_void_p
        L.t3do_destroy.argtypes = [C.c_void_p]
        L.t3do_mesh_register.argtypes = [C.c_void_p, C.c_uint64, C.POINTER(A.MeshDesc)]
        L.t3do_volume_register.argtypes = [C.c_void_p, C.c_uint64, FP, C.c_uint32]
        L.t3do_volume_set_cell_lights.argtypes = [C.c_void_p, C.c_uint64, C.POINTER(C.c_uint32), C.POINTER(C.c_uint32)]
        L.t3do_render.argtypes = [C.c_void_p, C.POINTER(A.RenderDesc), C.POINTER(OracleResult)]
        L.t3do_result_free.argtypes = [C.POINTER(OracleResult)]
        L.t3do_last_error.restype = C.c_char_p
        L.t3do_last_error.argtypes = [C.c_void_p]
        U8 = C.POINTER(C.c_uint8)
        L.t3do_texture_register.argtypes = [C.c_void_p, C.c_uint64, C.c_uint32, C.c_uint32, U8]
        L.t3do_rasterize.argtypes = [C.c_void_p, C.POINTER(A.RenderDesc), C.POINTER(OracleResult), U8, U8]
        for n in ("t3do_mat4_mult",):
            getattr(L, n).argtypes = [M4, M4, M4]
        L.t3do_bone_palettes.argtypes = [C.POINTER(A.BonePose), C.c_uint32, M4, M4]
        L.t3do_bone_palettes.restype = C.c_int
        L.t3do_animate_poses.argtypes = [C.POINTER(A.AnimChannel), C.c_uint32, C.c_float, C.POINTER(A.BonePose),
                                         C.c_uint32]
        L.t3do_animate_poses.restype = None
        L.t3do_animation_register.argtypes = [C.c_void_p, C.c_uint64, C.POINTER(A.AnimChannel), C.c_uint32]
        L.t3do_go_sin.argtypes = [C.c_double]
        L.t3do_go_sin.restype = C.c_double
        L.t3do_go_atan2.argtypes = [C.c_double, C.c_double]
        L.t3do_go_atan2.restype = C.c_double
        L.t3do_quat_slerp.argtypes = [FP, FP, C.c_float, FP]
        L.t3do_quat_slerp.restype = None
        L.t3do_quat_to_mat4.argtypes = [FP, M4]
        L.t3do_quat_to_mat4.restype = None
        L.t3do_mat4_inverted.argtypes = [M4, M4]
        L.t3do_mat4_transposed.argtypes = [M4, M4]
        L.t3do_mat4_decompose.argtypes = [M4, FP, FP, M4]
        L.t3do_mat4_mult_vec.argtypes = [M4, FP, FP]
        L.t3do_mat4_mult_vec_w.argtypes = [M4, FP, FP]
        L.t3do_projection_perspective.argtypes = [C.c_float] * 5 + [M4]
        L.t3do_view_matrix.argtypes = [C.POINTER(A.CameraDesc), M4]
        L.t3do_vp_matrix.argtypes = [C.POINTER(A.CameraDesc), M4]
        L.t3do_vec3_unit.argtypes = [FP, FP]
        L.t3do_sphere_in_frustum.argtypes = [C.POINTER(A.CameraDesc), FP, C.c_float]
        L.t3do_model_frustum_sphere.argtypes = [M4, C.c_int, FP, FP, FP, FP]
        L.t3do_inverted_cam_pos.argtypes = [M4, FP, FP]
        L.t3do_bucket_sort.argtypes = [FP, C.c_uint32, C.c_float, C.c_float, C.c_int32, C.POINTER(C.c_uint32)]
        L.t3do_bucket_sort.restype = C.c_uint32
        L.t3do_bucket_bin.argtypes = [C.c_float, C.c_float, C.c_float, C.c_int32]
        L.t3do_bucket_bin.restype = C.c_uint32
        L.t3do_go_stable_sort_table.argtypes = [C.POINTER(C.c_int32), C.c_uint32, C.POINTER(C.c_uint8)]
        _lib = L
    return _lib


def animate_poses(anim, time, poses):
    """Copy of `poses` after sampling `anim` at `time` (animation.go:63-160,
    902-983)."""
    out = (A.BonePose * len(poses))()
    C.memmove(out, poses, C.sizeof(A.BonePose) * len(poses))
    arr, keep = anim.channels_desc()
    lib().t3do_animate_poses(arr, len(anim.channels), time, out, len(poses))
    return out


def bone_palettes(poses, root=None):
    """Palettes of an armature given as poses (ctypes array of BonePose,
    parents first), as Node.Transform() computes them (node.go:572-608):
    (n, 4, 4) float32, or None when a parent is not an earlier bone."""
    n = len(poses)
    out = (A.Mat4 * max(1, n))()
    r = lib().t3do_bone_palettes(poses, n, C.byref(A.mat4(root)) if root is not None else None, out)
    if r != 0:
        return None
    return np.frombuffer(out, dtype=np.float32, count=16 * n).reshape(n, 4, 4).copy()


def _vertices(ptr, n):
    if n == 0:
        return np.zeros((0, 12), dtype=np.float32)
    return np.ctypeslib.as_array(C.cast(ptr, C.POINTER(C.c_float)), shape=(n * 12,)).reshape(n, 12).copy()


class Frame:
    """Outputs of one oracle render, pair-concatenated like the GPU download."""

    def __init__(self, res):
        n = int(res.vertex_count)
        self.pairs = [(p.model_index, p.part_index, p.triangle_count, p.survivor_count, p.vertex_offset)
                      for p in res.pairs[:res.pair_count]]
        self.color = _vertices(res.color, n)
        self.depth = _vertices(res.depth, n)
        self.indices = np.ctypeslib.as_array(res.indices, shape=(n,)).copy() if n else np.zeros(0, np.uint32)
        t = n // 3
        self.triangle_ids = np.ctypeslib.as_array(res.triangle_ids, shape=(t,)).copy() if t else np.zeros(0, np.uint32)
        self.depth_keys = np.ctypeslib.as_array(res.depth_keys, shape=(t,)).copy() if t else np.zeros(0, np.float32)
        self.normals = _vertices(res.normals, n) if res.normals else None   # Camera.RenderNormals only
        self.survivors = t


class Oracle:
    """An oracle context holding registered meshes and volumes."""

    def __init__(self):
        self.L = lib()
        self.ctx = self.L.t3do_create()
        self._handles = {}
        self._refs = []   # registered meshes stay alive so their ids are not reused
        self._own_handle = 1 << 40   # above any handle a device context hands out
        self._anims = {}             # registered animations (copied by the oracle), kept alive by id

    def close(self):
        if self.ctx:
            self.L.t3do_destroy(self.ctx)
            self.ctx = None

    def __del__(self):
        self.close()

    def register_mesh(self, mesh, handle):
        d = mesh.desc()
        rc = self.L.t3do_mesh_register(self.ctx, handle, C.byref(d))
        if rc != 0:
            raise RuntimeError("oracle mesh register: %s" % self.L.t3do_last_error(self.ctx).decode())
        self._handles[id(mesh)] = handle
        self._refs.append(mesh)

    def register_volume(self, vol, handle):
        cells = np.ascontiguousarray(vol.cells, dtype=np.float32)
        self.L.t3do_volume_register(self.ctx, handle, cells.ctypes.data_as(C.POINTER(C.c_float)), len(cells))
        first, index = vol.cell_light_lists()
        if first is not None:
            U = C.POINTER(C.c_uint32)
            self.L.t3do_volume_set_cell_lights(self.ctx, handle, first.ctypes.data_as(U), index.ctypes.data_as(U))

    def register_animation(self, anim, handle):
        arr, keep = anim.channels_desc()
        self.L.t3do_animation_register(self.ctx, handle, arr, len(anim.channels))
        self._anims[id(anim)] = anim

    def handle_of(self, mesh):
        return self._handles[id(mesh)]

    def register_texture(self, tex):
        """Register a Texture under its handle (one of the oracle's own when
        it has none yet)."""
        if id(tex) in self._handles:
            return
        if not tex.handle:
            self._own_handle += 1
            tex.handle = self._own_handle
        h, w = tex.rgba.shape[:2]
        rc = self.L.t3do_texture_register(self.ctx, tex.handle, w, h, tex.rgba.ctypes.data_as(C.POINTER(C.c_uint8)))
        if rc != 0:
            raise RuntimeError("oracle texture register: %s" % self.L.t3do_last_error(self.ctx).decode())
        self._handles[id(tex)] = tex.handle
        self._refs.append(tex)

    def rasterize(self, camera, scene, lights, models):
        """Render, then draw the flushes into the camera's result textures:
        (frame, colour (H, W, 4) uint8, depth (H, W, 4) uint8 or None)."""
        for m in with_batched(models):
            for p in (m.mesh.mesh_parts if m.mesh is not None else []):
                if p.material is not None and p.material.texture is not None:
                    self.register_texture(p.material.texture)
        inp = RenderInputs(camera, scene, lights, models, self.handle_of)
        res = OracleResult()
        rc = self.L.t3do_render(self.ctx, C.byref(inp.render), C.byref(res))
        if rc != 0:
            raise RuntimeError("oracle render (%d): %s" % (rc, self.L.t3do_last_error(self.ctx).decode()))
        try:
            frame = Frame(res)
            W, H = camera.width, camera.height
            color = np.zeros((H, W, 4), dtype=np.uint8)
            depth = np.zeros((H, W, 4), dtype=np.uint8) if camera.render_depth else None
            U8 = C.POINTER(C.c_uint8)
            rc = self.L.t3do_rasterize(self.ctx, C.byref(inp.render), C.byref(res), color.ctypes.data_as(U8),
                                       depth.ctypes.data_as(U8) if depth is not None else None)
            if rc != 0:
                raise RuntimeError("oracle rasterize (%d): %s" % (rc, self.L.t3do_last_error(self.ctx).decode()))
            return frame, color, depth
        finally:
            self.L.t3do_result_free(C.byref(res))

    def render(self, camera, scene, lights, models):
        inp = RenderInputs(camera, scene, lights, models, self.handle_of)
        res = OracleResult()
        rc = self.L.t3do_render(self.ctx, C.byref(inp.render), C.byref(res))
        if rc != 0:
            raise RuntimeError("oracle render (%d): %s" % (rc, self.L.t3do_last_error(self.ctx).decode()))
        try:
            return Frame(res)
        finally:
            self.L.t3do_result_free(C.byref(res))

    def render_seconds(self, camera, scene, lights, models):
        """Seconds the C restatement takes for one Camera.Render (descriptor
        building and the conversion of its outputs to numpy excluded)."""
        inp = RenderInputs(camera, scene, lights, models, self.handle_of)
        res = OracleResult()
        t0 = time.perf_counter()
        rc = self.L.t3do_render(self.ctx, C.byref(inp.render), C.byref(res))
        el = time.perf_counter() - t0
        self.L.t3do_result_free(C.byref(res))
        if rc != 0:
            raise RuntimeError("oracle render (%d): %s" % (rc, self.L.t3do_last_error(self.ctx).decode()))
        return el

    def render_workload(self, wl):
        """Register the workload's meshes/volumes (handles = the ones the GPU
        side assigned, stored on each object) and render it."""
        for m in wl.meshes:
            if id(m) not in self._handles:
                h = m._handles.get("gpu")
                if h is None:               # not uploaded to a device: a handle of the oracle's own
                    self._own_handle += 1
                    h = self._own_handle
                self.register_mesh(m, h)
        for v in wl.volumes:
            if v.handle == 0:
                self._own_handle += 1
                v.handle = self._own_handle
            self.register_volume(v, v.handle)
        for an in wl.animations:
            if id(an) not in self._anims:
                if an.handle == 0:
                    self._own_handle += 1
                    an.handle = self._own_handle
                self.register_animation(an, an.handle)
        return self.render(wl.camera, wl.scene, wl.lights, wl.models)
