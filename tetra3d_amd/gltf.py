"""glTF documents through the library's loader (t3d_gltf_*, the mesh and skin
reading of LoadGLTFData, gltf.go:560-840 and 1449-1505).

    doc = Document(open("scene.glb", "rb").read())
    doc.meshes[0]            # a scene.Mesh holding the decoded arrays
    doc.skins[0]             # (joint node indices, inverse bind matrices (n, 4, 4))
    doc.upload(ctx)          # every mesh into the context's resident pool

Parsing runs on the host and needs no GPU.
"""
import ctypes as C

import numpy as np

from tetra3d_amd import _abi as A
from tetra3d_amd import runtime as R
from tetra3d_amd.scene import Material, Mesh, MeshPart

f32 = np.float32


class GltfMeshInfo(C.Structure):
    _fields_ = [
        ("name", C.c_char * 64),
        ("vertex_count", C.c_uint32),
        ("triangle_count", C.c_uint32),
        ("part_count", C.c_uint32),
        ("shape_key_count", C.c_uint32),
        ("has_colors", C.c_int32),
        ("skinned", C.c_int32),
    ]


def _lib():
    L = R.lib()
    if not getattr(L, "_gltf_ready", False):
        vp, P = C.c_void_p, C.POINTER
        L.t3d_gltf_parse.argtypes = [C.c_void_p, C.c_uint64, P(vp), C.c_char_p, C.c_uint32]
        L.t3d_gltf_free.argtypes = [vp]
        L.t3d_gltf_free.restype = None
        L.t3d_gltf_mesh_count.argtypes = [vp]
        L.t3d_gltf_mesh_count.restype = C.c_uint32
        L.t3d_gltf_mesh.argtypes = [vp, C.c_uint32, P(GltfMeshInfo), P(A.MeshDesc), P(P(C.c_int32))]
        L.t3d_gltf_skin_count.argtypes = [vp]
        L.t3d_gltf_skin_count.restype = C.c_uint32
        L.t3d_gltf_skin.argtypes = [vp, C.c_uint32, P(C.c_uint32), P(P(C.c_int32)), P(P(A.Mat4))]
        L.t3d_gltf_upload.argtypes = [vp, vp, P(C.c_uint64)]
        L._gltf_ready = True
    return L


class GltfError(Exception):
    pass


def _arr(ptr, n, dtype, shape):
    if not ptr or n == 0:
        return None
    return np.ctypeslib.as_array(ptr, shape=(n,)).astype(dtype, copy=True).reshape(shape)


class Document:
    """A parsed glTF document: meshes (scene.Mesh with the decoded arrays,
    one MeshPart per primitive), their summaries and primitive materials,
    and the skins."""

    def __init__(self, data):
        L = _lib()
        self._buf = bytes(data)
        err = C.create_string_buffer(512)
        h = C.c_void_p()
        rc = L.t3d_gltf_parse(self._buf, len(self._buf), C.byref(h), err, 512)
        if rc != 0:
            raise GltfError(err.value.decode(errors="replace"))
        self._h = h
        self.infos, self.meshes, self.part_materials = [], [], []
        try:
            for i in range(L.t3d_gltf_mesh_count(h)):
                info, d = GltfMeshInfo(), A.MeshDesc()
                pm = C.POINTER(C.c_int32)()
                rc = L.t3d_gltf_mesh(h, i, C.byref(info), C.byref(d), C.byref(pm))
                if rc != 0:
                    raise GltfError("mesh %d: status %d" % (i, rc))
                self.infos.append(info)
                self.meshes.append(self._mesh(info, d))
                self.part_materials.append([pm[k] for k in range(d.part_count)])
            self.skins = []
            for i in range(L.t3d_gltf_skin_count(h)):
                n = C.c_uint32()
                jp = C.POINTER(C.c_int32)()
                mp = C.POINTER(A.Mat4)()
                rc = L.t3d_gltf_skin(h, i, C.byref(n), C.byref(jp), C.byref(mp))
                if rc != 0:
                    raise GltfError("skin %d: status %d" % (i, rc))
                joints = [jp[k] for k in range(n.value)]
                inv = np.array([[[mp[k].m[r][c] for c in range(4)] for r in range(4)] for k in range(n.value)],
                               dtype=f32).reshape(-1, 4, 4)
                self.skins.append((joints, inv))
        except Exception:
            self.close()
            raise

    @staticmethod
    def _mesh(info, d):
        V, T, K = d.vertex_count, d.triangle_count, d.shape_key_count
        pos = _arr(d.positions, 3 * V, f32, (V, 3))
        nrm = _arr(d.normals, 3 * V, f32, (V, 3))
        idx = _arr(d.indices, 3 * T, np.uint32, (T, 3))
        col = _arr(d.colors, 4 * V, f32, (V, 4))
        uv = _arr(d.uvs, 2 * V, f32, (V, 2))
        bones = _arr(d.bone_indices, 4 * V, np.uint16, (V, 4))
        weights = _arr(d.bone_weights, 4 * V, f32, (V, 4))
        dims = (np.array(d.dimensions_min[:], dtype=f32), np.array(d.dimensions_max[:], dtype=f32))
        m = Mesh(info.name.decode(errors="replace"), pos, nrm, idx, colors=col, uvs=uv, bone_indices=bones,
                 bone_weights=weights, parts=[], dimensions=dims)
        m.mesh_parts = [MeshPart(m, Material("%s.%d" % (m.name, k)), d.parts[k].triangle_start, d.parts[k].triangle_count)
                        for k in range(d.part_count)]
        if K:
            m.shape_key_positions = _arr(d.shape_key_positions, 3 * K * V, f32, (K, V, 3))
            m.shape_key_normals = _arr(d.shape_key_normals, 3 * K * V, f32, (K, V, 3))
        return m

    def upload(self, ctx):
        """Every mesh into ctx's resident pool through t3d_gltf_upload; each
        scene.Mesh of this document then renders from its handle."""
        n = len(self.meshes)
        handles = (C.c_uint64 * max(1, n))()
        ctx._check(_lib().t3d_gltf_upload(ctx.ctx, self._h, handles))
        for m, h in zip(self.meshes, handles):
            ctx._mesh_handles[id(m)] = h
            ctx._mesh_refs[id(m)] = m
            m._handles["gpu"] = h
        return [handles[i] for i in range(n)]

    def close(self):
        if getattr(self, "_h", None):
            _lib().t3d_gltf_free(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
