"""ctypes mirror of include/t3d.h (the C-ABI of the geometry pipeline).

Field order and types must match the header exactly; tests check the struct
sizes against the library's own sizeof values.
"""
import ctypes as C

T3D_ABI_VERSION = 2

# status codes
T3D_OK = 0
T3D_ERR_INVALID_ARGUMENT = -1
T3D_ERR_OUT_OF_MEMORY = -2
T3D_ERR_BUDGET = -3
T3D_ERR_DEVICE = -4
T3D_ERR_UNSUPPORTED = -5
T3D_ERR_NOT_FOUND = -6
T3D_ERR_BUSY = -7
T3D_ERR_NOT_READY = -8
T3D_ERR_CANCELLED = -9

STATUS_NAMES = {
    0: "OK", -1: "INVALID_ARGUMENT", -2: "OUT_OF_MEMORY", -3: "BUDGET", -4: "DEVICE",
    -5: "UNSUPPORTED", -6: "NOT_FOUND", -7: "BUSY", -8: "NOT_READY", -9: "CANCELLED",
}

# TriangleSortMode (material.go:9-13)
SORT_BACK_TO_FRONT, SORT_FRONT_TO_BACK, SORT_NONE = 0, 1, 2
# TransparencyMode (material.go:15-27)
TRANSPARENCY_AUTO, TRANSPARENCY_OPAQUE, TRANSPARENCY_ALPHA_CLIP, TRANSPARENCY_TRANSPARENT = 0, 1, 2, 3
# LightingMode (material.go:35-38)
LIGHTING_DEFAULT, LIGHTING_FIXED_NORMALS, LIGHTING_DOUBLE_SIDED = 0, 1, 2
# LightVolumeShadingMode (material.go:50-55)
LV_PER_VERTEX_WITHOUT_NORMAL, LV_PER_VERTEX_WITH_NORMAL, LV_PER_OBJECT, LV_OFF = 0, 1, 2, 3
# BillboardUpDirection (material.go:29-32), BillboardedDepthMode (material.go:44-47)
BILLBOARD_UP_GLOBAL_Y, BILLBOARD_UP_CAMERA_Y = 0, 1
DEPTH_MODE_DEFAULT, DEPTH_MODE_UNBILLBOARDED = 0, 1
# FogBlendMode (world.go:8-11)
FOG_ADD, FOG_SUB, FOG_OVERWRITE, FOG_TRANSPARENT = 0, 1, 2, 3
FOG_CURVE_LINEAR, FOG_CURVE_OUT_CIRC, FOG_CURVE_IN_CIRC = 0, 1, 2
FILTER_NEAREST, FILTER_LINEAR = 0, 1
# light types
LIGHT_AMBIENT, LIGHT_POINT, LIGHT_DIRECTIONAL, LIGHT_VOLUME = 0, 1, 2, 3
# execution forms
EXEC_PLAIN, EXEC_BATCHED = 0, 1

F3 = C.c_float * 3
F4 = C.c_float * 4
F2 = C.c_float * 2
I3 = C.c_int32 * 3


class Mat4(C.Structure):
    _fields_ = [("m", (C.c_float * 4) * 4)]


class Vertex(C.Structure):
    _fields_ = [(n, C.c_float) for n in (
        "dst_x", "dst_y", "src_x", "src_y", "color_r", "color_g", "color_b", "color_a",
        "custom0", "custom1", "custom2", "custom3")]


class PartRange(C.Structure):
    _fields_ = [("triangle_start", C.c_uint32), ("triangle_count", C.c_uint32)]


class MeshDesc(C.Structure):
    _fields_ = [
        ("vertex_count", C.c_uint32),
        ("positions", C.POINTER(C.c_float)),
        ("normals", C.POINTER(C.c_float)),
        ("colors", C.POINTER(C.c_float)),
        ("uvs", C.POINTER(C.c_float)),
        ("bone_indices", C.POINTER(C.c_uint16)),
        ("bone_weights", C.POINTER(C.c_float)),
        ("shape_key_count", C.c_uint32),
        ("shape_key_positions", C.POINTER(C.c_float)),
        ("shape_key_normals", C.POINTER(C.c_float)),
        ("triangle_count", C.c_uint32),
        ("indices", C.POINTER(C.c_uint32)),
        ("part_count", C.c_uint32),
        ("parts", C.POINTER(PartRange)),
        ("dimensions_min", F3),
        ("dimensions_max", F3),
    ]


class MaterialDesc(C.Structure):
    _fields_ = [
        ("present", C.c_int32),
        ("visible", C.c_int32),
        ("color", F4),
        ("backface_culling", C.c_int32),
        ("triangle_sort_mode", C.c_int32),
        ("shadeless", C.c_int32),
        ("fogless", C.c_int32),
        ("lighting_mode", C.c_int32),
        ("light_volume_shading_mode", C.c_int32),
        ("transparency_mode", C.c_int32),
        ("blend_source_over", C.c_int32),
        ("render_order", C.c_int32),
        ("use_texture", C.c_int32),
        ("texture_size", F2),
        ("billboard", C.c_int32),
        ("custom_depth_function", C.c_int32),
        ("depth_mode", C.c_int32),
        ("billboard_up", C.c_int32),
        ("billboard_lock", C.c_int32 * 3),
        ("texture", C.c_uint64),
        ("texture_filter", C.c_int32),
        ("pad_texture", C.c_int32),
    ]


class LightDesc(C.Structure):
    pass


LightDesc._fields_ = [
    ("type", C.c_int32),
    ("visible", C.c_int32),
    ("color", F3),
    ("energy", C.c_float),
    ("range", C.c_float),
    ("transform", Mat4),
    ("world_rotation", Mat4),
    ("volume", C.c_uint64),
    ("volume_min", F3),
    ("volume_max", F3),
    ("cell_size", F3),
    ("cell_count", I3),
    ("offset", F3),
    ("light_bias", F3),
    ("use_lights", C.POINTER(LightDesc)),
    ("use_light_count", C.c_uint32),
    ("pad_use", C.c_int32),
]


class BonePose(C.Structure):
    _fields_ = [
        ("position", C.c_float * 3),
        ("scale", C.c_float * 3),
        ("parent", C.c_int32),
        ("pad", C.c_uint32),
        ("rotation", Mat4),
        ("inverse_bind", Mat4),
    ]


INTERPOLATION_LINEAR, INTERPOLATION_CONSTANT, INTERPOLATION_CUBIC = 0, 1, 2


class AnimTrack(C.Structure):
    _fields_ = [
        ("keyframe_count", C.c_uint32),
        ("interpolation", C.c_int32),
        ("times", C.POINTER(C.c_float)),
        ("values", C.POINTER(C.c_float)),
    ]


class AnimChannel(C.Structure):
    _fields_ = [("position", AnimTrack), ("scale", AnimTrack), ("rotation", AnimTrack)]


class ModelDesc(C.Structure):
    _fields_ = [
        ("mesh", C.c_uint64),
        ("transform", Mat4),
        ("world_rotation", Mat4),
        ("color", F4),
        ("visible", C.c_int32),
        ("shadeless", C.c_int32),
        ("frustum_culling", C.c_int32),
        ("auto_batch_static", C.c_int32),
        ("skinned", C.c_int32),
        ("bone_count", C.c_uint32),
        ("bone_influences", C.POINTER(Mat4)),
        ("shape_key_weights", C.POINTER(C.c_float)),
        ("materials", C.POINTER(MaterialDesc)),
        ("light_group_active", C.c_int32),
        ("light_group_count", C.c_uint32),
        ("light_group", C.POINTER(LightDesc)),
        ("has_vertex_transform_function", C.c_int32),
        ("has_vertex_clip_function", C.c_int32),
        ("bone_poses", C.POINTER(BonePose)),
        ("armature_parent", C.POINTER(Mat4)),
        ("animation", C.c_uint64),
        ("animation_time", C.c_float),
        ("pad_animation", C.c_uint32),
        ("dynamic_batched", C.c_int32),
        ("batch_owner", C.c_uint32),
        ("batch_part", C.c_uint32),
        ("batch_order", C.c_uint32),
    ]


class CameraDesc(C.Structure):
    _fields_ = [
        ("width", C.c_int32),
        ("height", C.c_int32),
        ("perspective", C.c_int32),
        ("fov", C.c_float),
        ("near", C.c_float),
        ("far", C.c_float),
        ("ortho_scale", C.c_float),
        ("transform", Mat4),
        ("world_rotation", Mat4),
        ("render_depth", C.c_int32),
        ("render_normals", C.c_int32),
        ("perspective_corrected_texture_mapping", C.c_int32),
        ("depth_margin", C.c_float),
        ("vertex_snapping", C.c_float),
        ("max_light_count", C.c_int32),
        ("has_max_light_count_focus", C.c_int32),
        ("max_light_count_focus", F3),
    ]


class WorldDesc(C.Structure):
    _fields_ = [
        ("present", C.c_int32),
        ("lighting_on", C.c_int32),
        ("fog_on", C.c_int32),
        ("fog_mode", C.c_int32),
        ("fog_color", F4),
        ("fog_range", F2),
        ("ambient_color", F3),
        ("ambient_energy", C.c_float),
        ("ambient_visible", C.c_int32),
        ("clear_color", F4),
        ("fog_curve", C.c_int32),
        ("dithered_fog_size", C.c_float),
    ]


class RenderDesc(C.Structure):
    _fields_ = [
        ("camera", C.POINTER(CameraDesc)),
        ("world", C.POINTER(WorldDesc)),
        ("light_count", C.c_uint32),
        ("lights", C.POINTER(LightDesc)),
        ("model_count", C.c_uint32),
        ("models", C.POINTER(ModelDesc)),
        ("exec_mode", C.c_int32),
        ("process_all_triangles", C.c_int32),
    ]


class PairResult(C.Structure):
    _fields_ = [
        ("model_index", C.c_uint32),
        ("part_index", C.c_uint32),
        ("triangle_count", C.c_uint32),
        ("survivor_count", C.c_uint32),
        ("vertex_offset", C.c_uint64),
    ]


class KernelTime(C.Structure):
    _fields_ = [
        ("name", C.c_char * 32),
        ("total_ms", C.c_double),
        ("launches", C.c_uint32),
        ("algorithmic_bytes", C.c_uint64),
    ]


class Tuning(C.Structure):
    _fields_ = [
        ("vertex_block", C.c_uint32),
        ("vertices_per_thread", C.c_uint32),
        ("triangle_tile", C.c_uint32),
        ("scatter_waves", C.c_uint32),
        ("use_graph", C.c_uint32),
        ("store_mode", C.c_uint32),
        ("assemble_split", C.c_uint32),
        ("xcd_swizzle", C.c_uint32),
        ("vertex_split", C.c_uint32),
        ("cull_block", C.c_uint32),
        ("model_stage", C.c_uint32),
    ]


def mat4(rows):
    """Build a Mat4 from a 4x4 nested sequence or numpy array (row-major)."""
    m = Mat4()
    for i in range(4):
        for j in range(4):
            m.m[i][j] = float(rows[i][j])
    return m


def mat4_to_list(m):
    return [[m.m[i][j] for j in range(4)] for i in range(4)]
