"""ctypes binding of libt3d.so (include/t3d.h) and the Camera.Render entry.

The library is the only compute path: if libt3d.so is missing or no HIP
device is present, the calls here raise instead of computing anything on the
CPU.
"""
import ctypes as C
import os
import re

import numpy as np

from . import _abi as A
from .scene import RenderInputs, with_batched

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libt3d.so")
HEADER = os.path.join(os.path.dirname(HERE), "include", "t3d.h")


class T3DError(RuntimeError):
    def __init__(self, status, msg):
        super().__init__("%s (%d): %s" % (A.STATUS_NAMES.get(status, "?"), status, msg))
        self.status = status


_lib = None


def header_symbols():
    """Entry points declared in include/t3d.h."""
    src = open(HEADER).read()
    return sorted(set(re.findall(r"\b(t3d_[a-z0-9_]+)\s*\(", src)))


def lib():
    """Load libt3d.so (raises if it has not been built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise RuntimeError("libt3d.so is not built; run `python -m tetra3d_amd.build` (the HIP path has no fallback)")
    L = C.CDLL(LIB_PATH)
    P = C.POINTER
    vp = C.c_void_p
    L.t3d_abi_version.restype = C.c_int32
    L.t3d_struct_sizes.argtypes = [P(C.c_uint32), C.c_uint32]
    L.t3d_struct_sizes.restype = C.c_uint32
    L.t3d_ctx_create.argtypes = [C.c_int32, C.c_uint64, P(vp)]
    L.t3d_ctx_destroy.argtypes = [vp]
    L.t3d_last_error.argtypes = [vp]
    L.t3d_last_error.restype = C.c_char_p
    L.t3d_ctx_memory_used.argtypes = [vp]
    L.t3d_ctx_memory_used.restype = C.c_uint64
    L.t3d_mesh_upload.argtypes = [vp, P(A.MeshDesc), P(C.c_uint64)]
    L.t3d_mesh_free.argtypes = [vp, C.c_uint64]
    L.t3d_volume_upload.argtypes = [vp, P(C.c_float), C.c_uint32, P(C.c_uint64)]
    L.t3d_volume_free.argtypes = [vp, C.c_uint64]
    L.t3d_volume_set_cell_lights.argtypes = [vp, C.c_uint64, P(C.c_uint32), P(C.c_uint32)]
    L.t3d_animation_upload.argtypes = [vp, P(A.AnimChannel), C.c_uint32, P(C.c_uint64)]
    L.t3d_animation_free.argtypes = [vp, C.c_uint64]
    L.t3d_frame_create.argtypes = [vp, P(vp)]
    L.t3d_frame_destroy.argtypes = [vp, vp]
    L.t3d_render.argtypes = [vp, vp, P(A.RenderDesc)]
    L.t3d_render_async.argtypes = [vp, vp, P(A.RenderDesc)]
    L.t3d_render_noalloc.argtypes = [vp, vp, P(A.RenderDesc), vp, vp, vp, C.c_uint64]
    L.t3d_debug_poison_outputs.argtypes = [vp, C.c_uint32]
    L.t3d_frame_wait.argtypes = [vp, vp]
    L.t3d_frame_pairs.argtypes = [vp, P(C.c_uint32), P(P(A.PairResult))]
    L.t3d_frame_device_outputs.argtypes = [vp, P(vp), P(vp), P(vp)]
    L.t3d_frame_download.argtypes = [vp, vp, P(A.Vertex), P(A.Vertex), P(C.c_uint32)]
    L.t3d_frame_download_triangles.argtypes = [vp, vp, P(C.c_uint32)]
    L.t3d_frame_device_normals.argtypes = [vp, P(vp)]
    L.t3d_frame_download_normals.argtypes = [vp, vp, P(A.Vertex)]
    L.t3d_texture_upload.argtypes = [vp, C.c_uint32, C.c_uint32, P(C.c_uint8), P(C.c_uint64)]
    L.t3d_texture_free.argtypes = [vp, C.c_uint64]
    L.t3d_frame_rasterize.argtypes = [vp, vp]
    L.t3d_frame_device_images.argtypes = [vp, P(vp), P(vp)]
    L.t3d_frame_download_images.argtypes = [vp, vp, P(C.c_uint8), P(C.c_uint8)]
    L.t3d_frame_last_ms.argtypes = [vp]
    L.t3d_frame_last_ms.restype = C.c_double
    L.t3d_profile_enable.argtypes = [vp, C.c_int32]
    L.t3d_profile_read.argtypes = [vp, P(A.KernelTime), C.c_uint32, P(C.c_uint32), C.c_int32]
    L.t3d_get_tuning.argtypes = [vp, P(A.Tuning)]
    L.t3d_set_tuning.argtypes = [vp, P(A.Tuning)]
    L.t3d_autotune.argtypes = [vp, vp, P(A.RenderDesc), C.c_uint32, P(A.Tuning)]
    L.t3d_selftest_fp.argtypes = [vp, P(C.c_float), P(C.c_float), P(C.c_float), C.c_uint32]
    L.t3d_selftest_copy.argtypes = [vp, C.c_uint64, C.c_uint32, P(C.c_double)]
    L.t3d_debug_model_spheres.argtypes = [P(A.Mat4), P(C.c_float), P(C.c_uint8), C.c_uint32, P(C.c_float),
                                          P(C.c_float)]
    for name in ("t3d_ctx_create", "t3d_ctx_destroy", "t3d_mesh_upload", "t3d_mesh_free", "t3d_volume_upload",
                 "t3d_volume_free", "t3d_volume_set_cell_lights", "t3d_frame_create", "t3d_frame_destroy", "t3d_render", "t3d_frame_pairs",
                 "t3d_frame_device_outputs", "t3d_frame_download", "t3d_frame_download_triangles",
                 "t3d_profile_enable", "t3d_profile_read", "t3d_get_tuning", "t3d_set_tuning", "t3d_autotune",
                 "t3d_selftest_fp", "t3d_selftest_copy", "t3d_debug_model_spheres", "t3d_animation_upload",
                 "t3d_animation_free", "t3d_render_noalloc", "t3d_frame_device_normals",
                 "t3d_frame_download_normals", "t3d_texture_upload", "t3d_texture_free", "t3d_frame_rasterize",
                 "t3d_frame_device_images", "t3d_frame_download_images"):
        getattr(L, name).restype = C.c_int32
    _lib = L
    return L


def struct_sizes():
    L = lib()
    buf = (C.c_uint32 * 32)()
    n = L.t3d_struct_sizes(buf, 32)
    return list(buf[:n])


class Frame:
    """Host copy of one render's outputs, pair-concatenated in draw order."""

    def __init__(self, pairs, color, depth, indices, triangle_ids, ms, normals=None):
        self.pairs = pairs            # (model, part, triangles, survivors, device vertex offset)
        self.color = color            # (3*S, 12) float32 ebiten vertices
        self.depth = depth
        self.indices = indices        # (3*S,) uint32, pair-local
        self.triangle_ids = triangle_ids
        self.ms = ms
        self.normals = normals        # (3*S, 12) normal vertices when Camera.RenderNormals, else None

    @property
    def survivors(self):
        return int(sum(p[3] for p in self.pairs))


class Context:
    """A t3d_ctx on one HIP device: resident meshes, light volumes and frames."""

    def __init__(self, device=0, hbm_budget=0):
        self.L = lib()
        self.ctx = C.c_void_p()
        rc = self.L.t3d_ctx_create(device, hbm_budget, C.byref(self.ctx))
        if rc != 0:
            raise T3DError(rc, "t3d_ctx_create failed (is a HIP device visible?)")
        self.frame = C.c_void_p()
        self._check(self.L.t3d_frame_create(self.ctx, C.byref(self.frame)))
        self._extra_frames = []
        self._mesh_handles = {}   # id(mesh) -> handle
        self._mesh_refs = {}      # id(mesh) -> mesh: keeps ids from being reused while uploaded
        self._tex_refs = {}       # id(texture) -> texture, uploaded
        self._image_size = {}     # frame -> (width, height, render depth) of its last render

    def _check(self, rc):
        if rc != 0:
            raise T3DError(rc, self.L.t3d_last_error(self.ctx).decode(errors="replace"))
        return rc

    def close(self):
        if self.ctx:
            for f in self._extra_frames:
                self.L.t3d_frame_destroy(self.ctx, f)
            self._extra_frames = []
            if self.frame:
                self.L.t3d_frame_destroy(self.ctx, self.frame)
                self.frame = None
            self.L.t3d_ctx_destroy(self.ctx)
            self.ctx = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    @property
    def memory_used(self):
        return int(self.L.t3d_ctx_memory_used(self.ctx))

    # -- resources ---------------------------------------------------------
    def upload_mesh(self, mesh):
        h = self._mesh_handles.get(id(mesh))
        if h is not None:
            return h
        d = mesh.desc()
        out = C.c_uint64()
        self._check(self.L.t3d_mesh_upload(self.ctx, C.byref(d), C.byref(out)))
        self._mesh_handles[id(mesh)] = out.value
        self._mesh_refs[id(mesh)] = mesh
        mesh._handles["gpu"] = out.value
        return out.value

    def free_mesh(self, mesh):
        h = self._mesh_handles[id(mesh)]
        self._check(self.L.t3d_mesh_free(self.ctx, h))   # BUSY while a server job references it
        del self._mesh_handles[id(mesh)]
        self._mesh_refs.pop(id(mesh), None)
        mesh._handles.pop("gpu", None)

    def upload_volume(self, vol):
        cells = np.ascontiguousarray(vol.cells, dtype=np.float32)
        out = C.c_uint64()
        self._check(self.L.t3d_volume_upload(self.ctx, cells.ctypes.data_as(C.POINTER(C.c_float)), len(cells),
                                             C.byref(out)))
        vol.handle = out.value
        vol._lists_synced = None
        self.sync_cell_lights(vol)
        return out.value

    def sync_cell_lights(self, vol):
        """The volume's cell light lists to the device when they changed since
        the last upload (t3d_volume_set_cell_lights)."""
        if getattr(vol, "_lists_synced", None) == vol.lists_version:
            return
        first, index = vol.cell_light_lists()
        U = C.POINTER(C.c_uint32)
        self._check(self.L.t3d_volume_set_cell_lights(
            self.ctx, vol.handle, first.ctypes.data_as(U) if first is not None else None,
            index.ctypes.data_as(U) if index is not None else None))
        vol._lists_synced = vol.lists_version

    def upload_animation(self, anim):
        arr, keep = anim.channels_desc()
        out = C.c_uint64()
        self._check(self.L.t3d_animation_upload(self.ctx, arr, len(anim.channels), C.byref(out)))
        anim.handle = out.value
        return out.value

    def free_animation(self, anim):
        self._check(self.L.t3d_animation_free(self.ctx, anim.handle))
        anim.handle = 0

    def mesh_handle(self, mesh):
        return self._mesh_handles[id(mesh)]

    def prepare(self, workload):
        for m in workload.meshes:
            self.upload_mesh(m)
        self.upload_textures(workload.models)
        for v in workload.volumes:
            if not v.handle:
                self.upload_volume(v)
            else:
                self.sync_cell_lights(v)
        for an in workload.animations:
            if not an.handle:
                self.upload_animation(an)

    # -- tuning / profiling -------------------------------------------------
    def tuning(self):
        t = A.Tuning()
        self._check(self.L.t3d_get_tuning(self.ctx, C.byref(t)))
        return t

    def set_tuning(self, **kw):
        t = self.tuning()
        for k, v in kw.items():
            setattr(t, k, v)
        self._check(self.L.t3d_set_tuning(self.ctx, C.byref(t)))

    def profile(self, on=True):
        self._check(self.L.t3d_profile_enable(self.ctx, int(on)))

    def profile_read(self, reset=True):
        buf = (A.KernelTime * 32)()
        n = C.c_uint32()
        self._check(self.L.t3d_profile_read(self.ctx, buf, 32, C.byref(n), int(reset)))
        return {buf[i].name.decode(): (buf[i].total_ms, buf[i].launches, buf[i].algorithmic_bytes)
                for i in range(n.value)}

    # -- rendering ---------------------------------------------------------
    def inputs(self, camera, scene, lights, models, exec_mode=A.EXEC_BATCHED):
        return RenderInputs(camera, scene, lights, models, self.mesh_handle, exec_mode)

    def render_inputs(self, inp):
        """Run one Camera.Render on the device; returns nothing (outputs stay in HBM)."""
        self._check(self.L.t3d_render(self.ctx, self.frame, C.byref(inp.render)))
        self._image_size[id(self.frame)] = (inp.cam.width, inp.cam.height, bool(inp.cam.render_depth))

    # -- raster sink --------------------------------------------------------
    def upload_texture(self, tex):
        """Upload a Texture (t3d_texture_upload); its handle goes into material descriptors."""
        if id(tex) in self._tex_refs:
            return tex.handle
        h, w = tex.rgba.shape[:2]
        out = C.c_uint64()
        self._check(self.L.t3d_texture_upload(self.ctx, w, h, tex.rgba.ctypes.data_as(C.POINTER(C.c_uint8)),
                                              C.byref(out)))
        tex.handle = out.value
        self._tex_refs[id(tex)] = tex
        return tex.handle

    def upload_textures(self, models):
        for m in with_batched(models):
            for p in (m.mesh.mesh_parts if m.mesh is not None else []):
                if p.material is not None and p.material.texture is not None:
                    self.upload_texture(p.material.texture)

    def rasterize(self, frame=None):
        """Draw the frame's last render into its colour and depth images on the device."""
        self._check(self.L.t3d_frame_rasterize(self.ctx, frame or self.frame))

    def download_images(self, frame=None):
        """The rasterized images: colour (H, W, 4) uint8 premultiplied RGBA, and depth
        likewise or None without RenderDepth."""
        frame = frame or self.frame
        w, h, rdepth = self._image_size[id(frame)]
        color = np.zeros((h, w, 4), dtype=np.uint8)
        depth = np.zeros((h, w, 4), dtype=np.uint8) if rdepth else None
        U8 = C.POINTER(C.c_uint8)
        self._check(self.L.t3d_frame_download_images(self.ctx, frame, color.ctypes.data_as(U8),
                                                     depth.ctypes.data_as(U8) if depth is not None else None))
        return color, depth

    def render_noalloc(self, inp, color, depth, indices, capacity, frame=None):
        """Camera.Render into caller-owned device arrays (device addresses as
        ints, each `capacity` vertices long) with no device allocation
        (t3d_render_noalloc); results are read with pairs(frame)."""
        self._check(self.L.t3d_render_noalloc(self.ctx, frame or self.frame, C.byref(inp.render), color, depth,
                                              indices, capacity))

    def poison_outputs(self, word=0x7FC0DEAD):
        """Fill frame outputs with `word` before every render (0: off); tests
        use a NaN pattern so unwritten entries cannot pass as an earlier
        render's values."""
        self._check(self.L.t3d_debug_poison_outputs(self.ctx, word))

    def new_frame(self):
        """An extra frame (outputs + scratch) of this context, for asynchronous renders."""
        f = C.c_void_p()
        self._check(self.L.t3d_frame_create(self.ctx, C.byref(f)))
        self._extra_frames.append(f)
        return f

    def render_async(self, inp, frame=None):
        """Enqueue one Camera.Render into `frame` (default: the context's frame) and return."""
        self._check(self.L.t3d_render_async(self.ctx, frame or self.frame, C.byref(inp.render)))

    def wait(self, frame=None):
        self._check(self.L.t3d_frame_wait(self.ctx, frame or self.frame))

    def autotune(self, inp, reps=3):
        t = A.Tuning()
        self._check(self.L.t3d_autotune(self.ctx, self.frame, C.byref(inp.render), reps, C.byref(t)))
        return t

    def last_ms(self):
        return float(self.L.t3d_frame_last_ms(self.frame))

    def pairs(self, frame=None):
        n = C.c_uint32()
        ptr = C.POINTER(A.PairResult)()
        self._check(self.L.t3d_frame_pairs(frame or self.frame, C.byref(n), C.byref(ptr)))
        return [(p.model_index, p.part_index, p.triangle_count, p.survivor_count, p.vertex_offset)
                for p in ptr[:n.value]]

    def download(self, triangles=True, frame=None):
        frame = frame or self.frame
        pairs = self.pairs(frame)
        S = int(sum(p[3] for p in pairs))
        color = np.zeros((3 * S, 12), dtype=np.float32)
        depth = np.zeros((3 * S, 12), dtype=np.float32)
        idx = np.zeros(3 * S, dtype=np.uint32)
        VP = C.POINTER(A.Vertex)
        self._check(self.L.t3d_frame_download(self.ctx, frame, color.ctypes.data_as(VP), depth.ctypes.data_as(VP),
                                              idx.ctypes.data_as(C.POINTER(C.c_uint32))))
        tri = None
        if triangles:
            tri = np.zeros(S, dtype=np.uint32)
            self._check(self.L.t3d_frame_download_triangles(self.ctx, frame,
                                                            tri.ctypes.data_as(C.POINTER(C.c_uint32))))
        normals = None
        dev = C.c_void_p()
        if self.L.t3d_frame_device_normals(frame, C.byref(dev)) == 0:   # the render filled a normal list
            normals = np.zeros((3 * S, 12), dtype=np.float32)
            self._check(self.L.t3d_frame_download_normals(self.ctx, frame, normals.ctypes.data_as(VP)))
        return Frame(pairs, color, depth, idx, tri, self.last_ms(), normals)

    def render(self, camera, scene, lights, models, exec_mode=A.EXEC_BATCHED, download=True):
        """Camera.Render(scene, lights, models) (camera.go:1269) on the GPU."""
        for m in with_batched(models):
            if m.mesh is not None:
                self.upload_mesh(m.mesh)
        self.upload_textures(models)
        for l in lights:
            if getattr(l, "TYPE", None) == A.LIGHT_VOLUME:
                if not l.handle:
                    self.upload_volume(l)
                else:
                    self.sync_cell_lights(l)
        inp = self.inputs(camera, scene, lights, models, exec_mode)
        self.render_inputs(inp)
        return self.download() if download else None

    def render_workload(self, wl, exec_mode=A.EXEC_BATCHED, download=True):
        self.prepare(wl)
        return self.render(wl.camera, wl.scene, wl.lights, wl.models, exec_mode, download)

    def selftest_copy(self, nbytes, reps=20):
        ms = C.c_double()
        self._check(self.L.t3d_selftest_copy(self.ctx, nbytes, reps, C.byref(ms)))
        return ms.value

    def selftest_fp(self, x, y):
        x = np.ascontiguousarray(x, dtype=np.float32)
        y = np.ascontiguousarray(y, dtype=np.float32)
        n = len(x)
        out = np.zeros(4 * n, dtype=np.float32)
        FP = C.POINTER(C.c_float)
        self._check(self.L.t3d_selftest_fp(self.ctx, x.ctypes.data_as(FP), y.ctypes.data_as(FP),
                                           out.ctypes.data_as(FP), n))
        return out.reshape(4, n)


def model_spheres(transforms, dims, skinned):
    """Bounding spheres (centre xyz, radius) of models by the render's
    four-at-a-time path and by the one-model path (t3d_debug_model_spheres);
    host only, no device needed."""
    L = lib()
    t = np.ascontiguousarray(transforms, dtype=np.float32).reshape(-1, 4, 4)
    n = len(t)
    d = np.ascontiguousarray(dims, dtype=np.float32).reshape(n, 6)
    sk = np.ascontiguousarray(skinned, dtype=np.uint8).reshape(n)
    ob = np.zeros((n, 4), dtype=np.float32)
    os_ = np.zeros((n, 4), dtype=np.float32)
    FP = C.POINTER(C.c_float)
    st = L.t3d_debug_model_spheres(t.ctypes.data_as(C.POINTER(A.Mat4)), d.ctypes.data_as(FP),
                                   sk.ctypes.data_as(C.POINTER(C.c_uint8)), n, ob.ctypes.data_as(FP),
                                   os_.ctypes.data_as(FP))
    if st != 0:
        raise RuntimeError("t3d_debug_model_spheres failed: %d" % st)
    return ob, os_
