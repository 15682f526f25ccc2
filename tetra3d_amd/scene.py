"""Host-side mirror of the Tetra3D scene API that Camera.Render consumes.

Names follow the reference (Node, Model, Mesh, MeshPart, Material, Camera,
AmbientLight, PointLight, DirectionalLight, LightVolume, World, Scene) so a
caller of the reference finds the same objects.  The scene graph stays on the
host, as in the reference; Camera.render() turns it into the C-ABI
descriptors of include/t3d.h and runs the geometry path on the GPU.

Node.transform() restates node.go:572-608 (Scale·Rotation·Translate·parent)
in float32 so world matrices are bit-identical to what the Go host computes.
"""
import ctypes as C

import numpy as np

from . import _abi as A

f32 = np.float32


# ---------------------------------------------------------------------------
# float32 matrix helpers (matrix.go), exact operation order
# ---------------------------------------------------------------------------

def identity():
    return np.eye(4, dtype=f32)


def mat_mult(a, b):
    """Matrix4.Mult (matrix_none.go:6-30): ((a0·b0 + a1·b1) + a2·b2) + a3·b3 per element."""
    a = np.asarray(a, dtype=f32)
    b = np.asarray(b, dtype=f32)
    return (((a[:, 0:1] * b[0:1, :] + a[:, 1:2] * b[1:2, :]) + a[:, 2:3] * b[2:3, :]) + a[:, 3:4] * b[3:4, :]).astype(f32)


def mat_translate(x, y, z):
    m = identity()
    m[3, 0], m[3, 1], m[3, 2] = f32(x), f32(y), f32(z)
    return m


def mat_scale(x, y, z):
    m = identity()
    m[0, 0], m[1, 1], m[2, 2] = f32(x), f32(y), f32(z)
    return m


def quat_to_matrix(x, y, z, w):
    """Quaternion.ToMatrix4 (quaternion.go:135-183): m1.Mult(m2)."""
    x, y, z, w = f32(x), f32(y), f32(z), f32(w)
    m1 = np.array([[w, z, -y, x], [-z, w, x, y], [y, -x, w, z], [-x, -y, -z, w]], dtype=f32)
    m2 = np.array([[w, z, -y, -x], [-z, w, x, -y], [y, -x, w, -z], [x, y, z, w]], dtype=f32)
    return mat_mult(m1, m2)


def vec3_unit(v):
    """Vector3.Unit (vector.go:258-266): divide by the float32 magnitude unless it is < 1e-8 or 1."""
    v = np.asarray(v, dtype=f32)
    l = f32(np.sqrt(np.float64(v[0] * v[0] + v[1] * v[1] + v[2] * v[2])))
    if l < f32(1e-8) or l == f32(1):
        return v.copy()
    return np.array([v[0] / l, v[1] / l, v[2] / l], dtype=f32)


def rotation_from_forward(direction):
    """A rotation Matrix4 whose row 2 (Forward) is the given unit direction."""
    fwd = vec3_unit(direction)
    up = np.array([0, 1, 0], dtype=f32)
    if abs(float(np.dot(fwd, up))) > 0.99:
        up = np.array([1, 0, 0], dtype=f32)
    right = np.cross(up, fwd).astype(f32)
    right = vec3_unit(right)
    up2 = np.cross(fwd, right).astype(f32)
    m = identity()
    m[0, :3], m[1, :3], m[2, :3] = right, vec3_unit(up2), fwd
    return m


# ---------------------------------------------------------------------------
# scene graph
# ---------------------------------------------------------------------------

class Node:
    """Node (node.go:406-436): local position, scale, rotation and parent."""

    def __init__(self, name=""):
        self.name = name
        self.position = np.zeros(3, dtype=f32)
        self.scale = np.ones(3, dtype=f32)
        self.rotation = identity()
        self.parent = None
        self.visible = True

    def set_local_position(self, x, y, z):
        self.position = np.array([x, y, z], dtype=f32)

    def set_local_scale(self, x, y, z):
        self.scale = np.array([x, y, z], dtype=f32)

    def set_local_rotation(self, rot):
        self.rotation = np.asarray(rot, dtype=f32).copy()

    def transform(self):
        """Node.Transform (node.go:572-608): Scale·Rotation·Translate·parent."""
        t = mat_scale(*self.scale)
        t = mat_mult(t, self.rotation)
        t = mat_mult(t, mat_translate(*self.position))
        if self.parent is not None:
            t = mat_mult(t, self.parent.transform())
        return t

    def world_rotation(self):
        """Node.WorldRotation (node.go:916-938) for parent chains without rotation."""
        p = self.parent
        while p is not None:
            if not np.allclose(p.rotation, identity(), atol=1e-4):
                t = self.transform()
                r = identity()
                for i in range(3):
                    row = t[i, :3]
                    r[i, :3] = vec3_unit(row)
                    r[i, 3] = t[i, 3]
                return r
            p = p.parent
        return self.rotation.copy()

    def world_position(self):
        return self.transform()[3, :3].copy()


class Material:
    """Material (material.go:66-131) with NewMaterial defaults (material.go:148-172)."""

    def __init__(self, name=""):
        self.name = name
        self.color = np.array([1, 1, 1, 1], dtype=f32)
        self.backface_culling = True
        self.triangle_sort_mode = A.SORT_BACK_TO_FRONT
        self.shadeless = False
        self.fogless = False
        self.lighting_mode = A.LIGHTING_DEFAULT
        self.light_volume_shading_mode = A.LV_PER_VERTEX_WITHOUT_NORMAL
        self.transparency_mode = A.TRANSPARENCY_AUTO
        self.blend_source_over = True
        self.render_order = 0
        self.visible = True
        self.texture_size = None  # (w, h) when a texture is set and UseTexture is on
        self.texture = None       # a Texture, read by the raster sink (its handle goes in the descriptor)
        self.texture_filter = A.FILTER_NEAREST
        self.billboard = False
        self.billboard_up = A.BILLBOARD_UP_GLOBAL_Y
        self.billboard_lock = (False, False, False)
        self.custom_depth_function = False
        self.depth_mode = A.DEPTH_MODE_DEFAULT

    def desc(self):
        d = A.MaterialDesc()
        d.present = 1
        d.visible = int(self.visible)
        d.color = A.F4(*[float(c) for c in self.color])
        d.backface_culling = int(self.backface_culling)
        d.triangle_sort_mode = self.triangle_sort_mode
        d.shadeless = int(self.shadeless)
        d.fogless = int(self.fogless)
        d.lighting_mode = self.lighting_mode
        d.light_volume_shading_mode = self.light_volume_shading_mode
        d.transparency_mode = self.transparency_mode
        d.blend_source_over = int(self.blend_source_over)
        d.render_order = self.render_order
        if self.texture_size is not None:
            d.use_texture = 1
            d.texture_size = A.F2(float(self.texture_size[0]), float(self.texture_size[1]))
        if self.texture is not None:
            d.texture = self.texture.handle
        d.texture_filter = int(self.texture_filter)
        d.billboard = int(self.billboard)
        d.custom_depth_function = int(self.custom_depth_function)
        d.depth_mode = self.depth_mode
        d.billboard_up = self.billboard_up
        d.billboard_lock = (C.c_int32 * 3)(*[int(bool(x)) for x in self.billboard_lock])
        return d


class Texture:
    """An ebiten.Image used as a material texture: (h, w, 4) uint8 RGBA with
    premultiplied alpha, as ebiten stores images.  `handle` is set when it is
    uploaded (t3d_texture_upload) and names it in material descriptors."""

    def __init__(self, rgba):
        self.rgba = np.ascontiguousarray(rgba, dtype=np.uint8)
        assert self.rgba.ndim == 3 and self.rgba.shape[2] == 4
        self.handle = 0

    @property
    def size(self):
        return (self.rgba.shape[1], self.rgba.shape[0])


class MeshPart:
    """MeshPart (mesh.go:2183-2190): a triangle range drawn with one material."""

    def __init__(self, mesh, material, triangle_start, triangle_count):
        self.mesh = mesh
        self.material = material
        self.triangle_start = triangle_start
        self.triangle_count = triangle_count


class Mesh:
    """Mesh (mesh.go:260-302): SoA vertex arrays, triangles and parts."""

    def __init__(self, name, positions, normals, indices, colors=None, uvs=None,
                 bone_indices=None, bone_weights=None, parts=None, dimensions=None):
        self.name = name
        self.positions = np.ascontiguousarray(positions, dtype=f32).reshape(-1, 3)
        self.normals = np.ascontiguousarray(normals, dtype=f32).reshape(-1, 3)
        self.indices = np.ascontiguousarray(indices, dtype=np.uint32).reshape(-1, 3)
        self.colors = None if colors is None else np.ascontiguousarray(colors, dtype=f32).reshape(-1, 4)
        self.uvs = None if uvs is None else np.ascontiguousarray(uvs, dtype=f32).reshape(-1, 2)
        self.bone_indices = None if bone_indices is None else np.ascontiguousarray(bone_indices, dtype=np.uint16).reshape(-1, 4)
        self.bone_weights = None if bone_weights is None else np.ascontiguousarray(bone_weights, dtype=f32).reshape(-1, 4)
        self.shape_key_positions = None
        self.shape_key_normals = None
        if parts is None:
            parts = [MeshPart(self, Material(name), 0, len(self.indices))]
        self.mesh_parts = parts
        if dimensions is None:
            dimensions = (self.positions.min(axis=0), self.positions.max(axis=0))  # UpdateDimensions, mesh.go:729-764
        self.dimensions = (np.asarray(dimensions[0], dtype=f32), np.asarray(dimensions[1], dtype=f32))
        self._handles = {}

    @property
    def vertex_count(self):
        return len(self.positions)

    @property
    def triangle_count(self):
        return len(self.indices)

    def desc(self):
        """A t3d_mesh_desc pointing into this mesh's arrays (keep the mesh alive)."""
        d = A.MeshDesc()
        fp = C.POINTER(C.c_float)
        d.vertex_count = self.vertex_count
        d.positions = self.positions.ctypes.data_as(fp)
        d.normals = self.normals.ctypes.data_as(fp)
        if self.colors is not None:
            d.colors = self.colors.ctypes.data_as(fp)
        if self.uvs is not None:
            d.uvs = self.uvs.ctypes.data_as(fp)
        if self.bone_indices is not None:
            d.bone_indices = self.bone_indices.ctypes.data_as(C.POINTER(C.c_uint16))
            d.bone_weights = self.bone_weights.ctypes.data_as(fp)
        if self.shape_key_positions is not None:
            d.shape_key_count = len(self.shape_key_positions)
            d.shape_key_positions = self.shape_key_positions.ctypes.data_as(fp)
            d.shape_key_normals = self.shape_key_normals.ctypes.data_as(fp)
        d.triangle_count = self.triangle_count
        d.indices = self.indices.ctypes.data_as(C.POINTER(C.c_uint32))
        parts = (A.PartRange * len(self.mesh_parts))()
        for i, p in enumerate(self.mesh_parts):
            parts[i].triangle_start = p.triangle_start
            parts[i].triangle_count = p.triangle_count
        self._parts_keepalive = parts
        d.part_count = len(self.mesh_parts)
        d.parts = parts
        d.dimensions_min = A.F3(*[float(x) for x in self.dimensions[0]])
        d.dimensions_max = A.F3(*[float(x) for x in self.dimensions[1]])
        return d


class Model(Node):
    """Model (model.go:20-61) with NewModel defaults (model.go:66-90)."""

    def __init__(self, name, mesh):
        super().__init__(name)
        self.mesh = mesh
        self.color = np.array([1, 1, 1, 1], dtype=f32)
        self.shadeless = False
        self.frustum_culling = True
        self.auto_batch_static = False
        self.skinned = False
        self.bone_influences = None  # (B, 4, 4) float32: boneInfluence per bone
        self.bone_poses = None       # BonePose ctypes array: bones as node state, parents first (palette built on the device)
        self.armature_parent = None  # 4x4: Transform() of the root bones' parent node
        self.animation = None        # Animation sampled on the device for the posed bones
        self.animation_time = 0.0    # its playhead (AnimationPlayer.Playhead() when Update samples)
        self.shape_key_weights = None
        self.light_group = None      # list of lights when a LightGroup is active
        self.dynamic_batch_models = {}   # part index -> [Model], DynamicBatchModels (model.go:29)
        self.dynamic_batch_owner = None  # DynamicBatchOwner (model.go:30)

    def dynamic_batch_add(self, part_index, *models):
        """DynamicBatchAdd (model.go:217-247): the models draw inside this
        model's flush for its MeshPart `part_index`, and this model's own mesh
        is no longer drawn.  A model batched elsewhere moves here.  (The
        reference's MaxTriangleCount cap follows from 16-bit index lists; the
        GPU path writes 32-bit indices and has no such cap.)"""
        for other in models:
            if other is self or any(other is m for ms in self.dynamic_batch_models.values() for m in ms):
                continue
            if other.dynamic_batch_owner is not None:
                other.dynamic_batch_owner.dynamic_batch_remove(other)
            self.dynamic_batch_models.setdefault(part_index, []).append(other)
            other.dynamic_batch_owner = self

    def dynamic_batch_remove(self, *models):
        """DynamicBatchRemove (model.go:249-270)."""
        for m in models:
            for part in list(self.dynamic_batch_models):
                lst = self.dynamic_batch_models[part]
                for i, existing in enumerate(lst):
                    if existing is m:
                        del lst[i]
                        m.dynamic_batch_owner = None
                        break
        for part in [p for p, lst in self.dynamic_batch_models.items() if not lst]:
            del self.dynamic_batch_models[part]

    def desc(self, mesh_handle, keep):
        d = A.ModelDesc()
        d.mesh = mesh_handle
        d.transform = A.mat4(self.transform())
        d.world_rotation = A.mat4(self.world_rotation())
        d.color = A.F4(*[float(c) for c in self.color])
        d.visible = int(self.visible)
        d.shadeless = int(self.shadeless)
        d.frustum_culling = int(self.frustum_culling)
        d.auto_batch_static = int(self.auto_batch_static)
        d.skinned = int(self.skinned)
        if self.skinned and self.bone_poses is not None:
            keep.append(self.bone_poses)
            d.bone_count = len(self.bone_poses)
            d.bone_poses = C.cast(self.bone_poses, C.POINTER(A.BonePose))
            if self.armature_parent is not None:
                root = A.mat4(self.armature_parent)
                keep.append(root)
                d.armature_parent = C.pointer(root)
        elif self.skinned:
            bones = np.ascontiguousarray(self.bone_influences, dtype=f32).reshape(-1, 4, 4)
            keep.append(bones)
            d.bone_count = len(bones)
            d.bone_influences = bones.ctypes.data_as(C.POINTER(A.Mat4))
        if self.animation is not None:   # the library refuses one on a model without bone poses
            if not self.animation.handle:
                raise ValueError("animation %r has no handle: upload it first" % self.animation.name)
            d.animation = self.animation.handle
            d.animation_time = float(self.animation_time)
        if self.shape_key_weights is not None:
            w = np.ascontiguousarray(self.shape_key_weights, dtype=f32)
            keep.append(w)
            d.shape_key_weights = w.ctypes.data_as(C.POINTER(C.c_float))
        mats = (A.MaterialDesc * len(self.mesh.mesh_parts))()
        for i, p in enumerate(self.mesh.mesh_parts):
            if p.material is not None:
                mats[i] = p.material.desc()
        keep.append(mats)
        d.materials = mats
        if self.light_group is not None:
            lg = (A.LightDesc * max(1, len(self.light_group)))()
            for i, l in enumerate(self.light_group):
                lg[i] = l.desc()
            keep.append(lg)
            d.light_group_active = 1
            d.light_group_count = len(self.light_group)
            d.light_group = lg
        return d


class Animation:
    """Animation (animation.go:188-260) of an armature: channel b animates
    bone b with up to three tracks (AnimationTrack, animation.go:40-160),
    each (times, values, interpolation): times ascending, values (K, 3) for
    "position" and "scale" or (K, 4) quaternions x, y, z, w for "rotation"."""

    def __init__(self, name, channel_count):
        self.name = name
        self.channels = [dict() for _ in range(channel_count)]
        self.handle = 0

    def add_track(self, bone, kind, times, values, interpolation=A.INTERPOLATION_LINEAR):
        width = 4 if kind == "rotation" else 3
        t = np.ascontiguousarray(times, dtype=f32).reshape(-1)
        v = np.ascontiguousarray(values, dtype=f32).reshape(len(t), width)
        self.channels[bone][kind] = (t, v, int(interpolation))

    def channels_desc(self):
        """ctypes AnimChannel array (and the arrays it points into)."""
        arr = (A.AnimChannel * max(1, len(self.channels)))()
        keep = [arr]
        for b, ch in enumerate(self.channels):
            for kind, (t, v, ip) in ch.items():
                tr = getattr(arr[b], kind)
                tr.keyframe_count = len(t)
                tr.interpolation = ip
                tr.times = t.ctypes.data_as(C.POINTER(C.c_float))
                tr.values = v.ctypes.data_as(C.POINTER(C.c_float))
                keep += [t, v]
        return arr, keep


class _Light(Node):
    TYPE = -1

    def __init__(self, name, r, g, b, energy):
        super().__init__(name)
        self.color = np.array([r, g, b], dtype=f32)
        self.energy = f32(energy)

    def desc(self):
        d = A.LightDesc()
        d.type = self.TYPE
        d.visible = int(self.visible)
        d.color = A.F3(*[float(c) for c in self.color])
        d.energy = float(self.energy)
        d.transform = A.mat4(self.transform())
        d.world_rotation = A.mat4(self.world_rotation())
        return d


class AmbientLight(_Light):
    """AmbientLight (light.go:80-132)."""
    TYPE = A.LIGHT_AMBIENT


class PointLight(_Light):
    """PointLight (light.go:137-285); range 0 selects inverse-square falloff."""
    TYPE = A.LIGHT_POINT

    def __init__(self, name, r, g, b, energy):
        super().__init__(name, r, g, b, energy)
        self.range = f32(0)

    def desc(self):
        d = super().desc()
        d.range = float(self.range)
        return d


class DirectionalLight(_Light):
    """DirectionalLight (light.go:290-374): shines along its rotation's row 2."""
    TYPE = A.LIGHT_DIRECTIONAL


class LightVolume(_Light):
    """LightVolume (light.go:376-680): a grid of cell colours."""
    TYPE = A.LIGHT_VOLUME

    def __init__(self, name, dims_min, dims_max, cell_size):
        super().__init__(name, 1, 1, 1, 1)
        self.dims_min = np.asarray(dims_min, dtype=f32)
        self.dims_max = np.asarray(dims_max, dtype=f32)
        self.cell_size = np.asarray(cell_size, dtype=f32)
        self.cell_count = self._count_cells()
        n = int(np.prod(self.cell_count))
        self.cells = np.zeros((n, 4), dtype=f32)
        self.cells[:, 3] = 1
        self.offset = np.zeros(3, dtype=f32)
        self.light_bias = np.zeros(3, dtype=f32)
        self.handle = 0
        self.cell_lights = {}        # cell index -> LightVolumeCell.useLights (light.go:1381)
        self.use_lights_render_set = []   # useLightsRenderSet (light.go:392), first-added order
        self.lists_version = 0       # bumped by every change of the cell lists

    def cell_index(self, x, y, z):
        """getLightCell's index of cell (x, y, z) (light.go:1267-1272)."""
        cx, cy, _ = self.cell_count
        return x + y * cx + z * cx * cy

    def _cell(self, cell):
        i = self.cell_index(*cell) if isinstance(cell, (tuple, list)) else int(cell)
        if not 0 <= i < len(self.cells):
            raise IndexError("cell %r outside the volume's %d cells" % (cell, len(self.cells)))
        return i

    def set_use_lights(self, cell, *lights):
        """LightVolumeCell.SetUseLights (light.go:1508-1513): append to the
        cell's lights (repeats kept); each joins the volume's render set."""
        i = self._cell(cell)
        for l in lights:
            if not any(l is x for x in self.use_lights_render_set):
                self.use_lights_render_set.append(l)
        self.cell_lights.setdefault(i, []).extend(lights)
        self.lists_version += 1

    def clear_use_lights(self, cell):
        """LightVolumeCell.ClearUseLights (light.go:1516-1518); the render set keeps its lights."""
        self.cell_lights.pop(self._cell(cell), None)
        self.lists_version += 1

    def use_lights(self, cell):
        return list(self.cell_lights.get(self._cell(cell), []))

    def cell_light_lists(self):
        """(cell_first[n + 1], light_index) as t3d_volume_set_cell_lights takes
        them, indices into use_lights_render_set; (None, None) when no cell
        names a light."""
        if not any(self.cell_lights.values()):
            return None, None
        pos = {id(l): k for k, l in enumerate(self.use_lights_render_set)}
        n = len(self.cells)
        first = np.zeros(n + 1, dtype=np.uint32)
        index = []
        for i in range(n):
            index += [pos[id(l)] for l in self.cell_lights.get(i, [])]
            first[i + 1] = len(index)
        return first, np.array(index if index else [0], dtype=np.uint32)

    def _count_cells(self):
        """Cell counts as LightVolumeResize's float32 loops produce them (light.go:467-494)."""
        counts = []
        for a in range(3):
            cs = f32(self.cell_size[a])
            half = f32(cs / f32(2))
            v = f32(self.dims_min[a] + half)
            hi = f32(self.dims_max[a] - half)
            n = 0
            while v <= hi:
                n += 1
                v = f32(v + cs)
            counts.append(n)
        return tuple(counts)

    def desc(self):
        d = super().desc()
        d.volume = self.handle
        d.volume_min = A.F3(*[float(x) for x in self.dims_min])
        d.volume_max = A.F3(*[float(x) for x in self.dims_max])
        d.cell_size = A.F3(*[float(x) for x in self.cell_size])
        d.cell_count = A.I3(*self.cell_count)
        d.offset = A.F3(*[float(x) for x in self.offset])
        d.light_bias = A.F3(*[float(x) for x in self.light_bias])
        if self.use_lights_render_set:
            arr = (A.LightDesc * len(self.use_lights_render_set))()
            for k, l in enumerate(self.use_lights_render_set):
                arr[k] = l.desc()
            self._use_keep = arr
            d.use_lights = arr
            d.use_light_count = len(self.use_lights_render_set)
        return d


class World:
    """World (world.go:25-63); the reference's default ambient energy is 0."""

    def __init__(self, name="world"):
        self.name = name
        self.lighting_on = True
        self.fog_on = True
        self.fog_mode = A.FOG_ADD
        self.fog_color = np.zeros(4, dtype=f32)
        self.fog_range = np.array([0, 1], dtype=f32)
        self.ambient_light = AmbientLight("ambient light", 1, 1, 1, 0)
        self.clear_color = np.array([0.08, 0.09, 0.1, 1], dtype=f32)   # NewWorld's ClearColor (world.go:56)
        self.fog_curve = A.FOG_CURVE_LINEAR
        self.dithered_fog_size = 0.0

    def desc(self):
        d = A.WorldDesc()
        d.present = 1
        d.lighting_on = int(self.lighting_on)
        d.fog_on = int(self.fog_on)
        d.fog_mode = self.fog_mode
        d.fog_color = A.F4(*[float(x) for x in self.fog_color])
        d.fog_range = A.F2(*[float(x) for x in self.fog_range])
        a = self.ambient_light
        d.ambient_color = A.F3(*[float(x) for x in a.color])
        d.ambient_energy = float(a.energy)
        d.ambient_visible = int(a.visible)
        d.clear_color = A.F4(*[float(x) for x in self.clear_color])
        d.fog_curve = int(self.fog_curve)
        d.dithered_fog_size = float(self.dithered_fog_size)
        return d


class Scene:
    """Scene (scene.go:12-31): a World plus the node lists handed to Render."""

    def __init__(self, name="scene", world=None):
        self.name = name
        self.world = world


class Camera(Node):
    """Camera (camera.go:256-321) with NewCamera defaults (camera.go:324-547)."""

    def __init__(self, name, w, h):
        super().__init__(name)
        self.width, self.height = int(w), int(h)
        self.perspective = True
        self.fov = f32(60)
        self.near = f32(0.1)
        self.far = f32(100)
        self.ortho_scale = f32(20)
        self.render_depth = True
        self.render_normals = False
        self.perspective_corrected_texture_mapping = False
        self.depth_margin = f32(0.04)
        self.vertex_snapping = f32(0)
        self.max_light_count = 0
        self.max_light_count_focus = None
        # not Camera state: ProcessVertices' processOnlyVisible for every render pair of the call (true in
        # Camera.Render, false for BakeLighting and the debug views)
        self.process_only_visible = True

    def desc(self):
        d = A.CameraDesc()
        d.width, d.height = self.width, self.height
        d.perspective = int(self.perspective)
        d.fov = float(self.fov)
        d.near = float(self.near)
        d.far = float(self.far)
        d.ortho_scale = float(self.ortho_scale)
        d.transform = A.mat4(self.transform())
        d.world_rotation = A.mat4(self.world_rotation())
        d.render_depth = int(self.render_depth)
        d.render_normals = int(self.render_normals)
        d.perspective_corrected_texture_mapping = int(self.perspective_corrected_texture_mapping)
        d.depth_margin = float(self.depth_margin)
        d.vertex_snapping = float(self.vertex_snapping)
        d.max_light_count = int(self.max_light_count)
        if self.max_light_count_focus is not None:
            d.has_max_light_count_focus = 1
            d.max_light_count_focus = A.F3(*[float(x) for x in self.max_light_count_focus])
        return d


def with_batched(models):
    """The models of a render call followed by every dynamically batched
    model of a batcher among them that is not listed itself: those draw
    through their owner's flush, so the call describes them too."""
    models = list(models)
    seen = {id(m) for m in models}
    k = 0
    while k < len(models):
        for lst in models[k].dynamic_batch_models.values():
            for b in lst:
                if id(b) not in seen:
                    seen.add(id(b))
                    models.append(b)
        k += 1
    return models


class RenderInputs:
    """The descriptors of one Camera.Render(scene, lights, models) call.

    `mesh_handle` maps a Mesh to the handle the executing side registered it
    under; `volume_handle` does the same for LightVolumes.  The object keeps
    every array the descriptors point into alive.
    """

    def __init__(self, camera, scene, lights, models, mesh_handle, exec_mode=A.EXEC_BATCHED):
        self.keep = []
        self.cam = camera.desc()
        self.world = scene.world.desc() if (scene is not None and scene.world is not None) else A.WorldDesc()
        self.lights = (A.LightDesc * max(1, len(lights)))()
        for i, l in enumerate(lights):
            self.lights[i] = l.desc()
        models = with_batched(models)
        index = {id(m): i for i, m in enumerate(models)}
        self.model_list = models
        self.models = (A.ModelDesc * max(1, len(models)))()
        for i, m in enumerate(models):
            h = mesh_handle(m.mesh) if m.mesh is not None else 0
            self.models[i] = m.desc(h, self.keep)
            if m.dynamic_batch_owner is not None and id(m.dynamic_batch_owner) not in index:
                # batched under a model this call does not render: never drawn (camera.go:1346)
                self.models[i].visible = 0
        for i, m in enumerate(models):
            for part, lst in m.dynamic_batch_models.items():
                for order, b in enumerate(lst):
                    d = self.models[index[id(b)]]
                    d.dynamic_batched = 1
                    d.batch_owner = i
                    d.batch_part = part
                    d.batch_order = order
        self.render = A.RenderDesc()
        self.render.camera = C.pointer(self.cam)
        self.render.world = C.pointer(self.world)
        self.render.light_count = len(lights)
        self.render.lights = self.lights
        self.render.model_count = len(models)
        self.render.models = self.models
        self.render.exec_mode = exec_mode
        self.render.process_all_triangles = 0 if getattr(camera, "process_only_visible", True) else 1
