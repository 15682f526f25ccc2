/*
 * t3d.h — C-ABI of the MI355X geometry pipeline for Tetra3D's Camera.Render.
 *
 * This boundary replaces the CPU path that Camera.Render runs per frame
 * (camera.go:1269-2114): per-model uniforms and frustum test, Model.ProcessVertices
 * (model.go:600-995: transform, skinning, cull, depth key), the 512-bin triangle
 * bucket sort (sortingTriangles.go:38-125), per-vertex lighting
 * (camera.go:1610-1644, light.go) and the ebiten.Vertex assembly
 * (camera.go:1646-1820).  It stops where the reference calls
 * DrawTriangles / DrawTrianglesShader (camera.go:1920-2036): the output is, per
 * render pair, the colour/depth vertex arrays and index list that flush() hands
 * to ebiten.
 *
 * Conventions
 *   - Matrices are Tetra3D Matrix4: row-major float32 [4][4], row vectors,
 *     v' = v·M, translation in row 3 (matrix.go:9-10, 619-637).
 *   - All arithmetic the library does is IEEE float32 in the reference's
 *     operation order, with no fused multiply-add.
 *   - Every entry point returns a t3d_status; on failure t3d_last_error(ctx)
 *     holds a message.  Nothing aborts across the boundary.
 *   - A context is externally synchronised: one calling thread at a time
 *     (the reference keeps all scratch in package globals, tetra3d.go:21-35,
 *     so it is single-threaded too).  The serving entry points are the
 *     exception and are documented as thread-safe below.
 *   - Handles are opaque non-zero uint64 values.
 */
#ifndef T3D_H
#define T3D_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define T3D_ABI_VERSION 2

typedef int32_t t3d_status;
enum {
  T3D_OK = 0,
  T3D_ERR_INVALID_ARGUMENT = -1, /* malformed descriptor (reference: panics, mesh.go:681, 2279) */
  T3D_ERR_OUT_OF_MEMORY = -2,    /* device allocation failed */
  T3D_ERR_BUDGET = -3,           /* request exceeds the context's HBM budget */
  T3D_ERR_DEVICE = -4,           /* HIP runtime error */
  T3D_ERR_UNSUPPORTED = -5,      /* feature outside the GPU contract (Go callbacks, billboards) */
  T3D_ERR_NOT_FOUND = -6,        /* unknown or freed handle */
  T3D_ERR_BUSY = -7,             /* resource still referenced by in-flight work */
  T3D_ERR_NOT_READY = -8,        /* job not finished yet (poll) */
  T3D_ERR_CANCELLED = -9         /* job was cancelled */
};

/* TriangleSortMode (material.go:9-13) */
enum { T3D_SORT_BACK_TO_FRONT = 0, T3D_SORT_FRONT_TO_BACK = 1, T3D_SORT_NONE = 2 };
/* TransparencyMode (material.go:15-27) */
enum {
  T3D_TRANSPARENCY_AUTO = 0,
  T3D_TRANSPARENCY_OPAQUE = 1,
  T3D_TRANSPARENCY_ALPHA_CLIP = 2,
  T3D_TRANSPARENCY_TRANSPARENT = 3
};
/* LightingMode (material.go:35-38) */
enum { T3D_LIGHTING_DEFAULT = 0, T3D_LIGHTING_FIXED_NORMALS = 1, T3D_LIGHTING_DOUBLE_SIDED = 2 };
/* LightVolumeShadingMode (material.go:50-55) */
enum {
  T3D_LV_PER_VERTEX_WITHOUT_NORMAL = 0,
  T3D_LV_PER_VERTEX_WITH_NORMAL = 1,
  T3D_LV_PER_OBJECT = 2,
  T3D_LV_OFF = 3
};
/* BillboardUpDirection (material.go:29-32) */
enum { T3D_BILLBOARD_UP_GLOBAL_Y = 0, T3D_BILLBOARD_UP_CAMERA_Y = 1 };
/* BillboardedDepthMode (material.go:44-47) */
enum { T3D_DEPTH_MODE_DEFAULT = 0, T3D_DEPTH_MODE_UNBILLBOARDED = 1 };
/* FogBlendMode (world.go:8-11) */
enum { T3D_FOG_ADD = 0, T3D_FOG_SUB = 1, T3D_FOG_OVERWRITE = 2, T3D_FOG_TRANSPARENT = 3 };
/* World.FogCurve (world.go:13-20) and Material.TextureFilterMode (material.go) */
enum { T3D_FOG_CURVE_LINEAR = 0, T3D_FOG_CURVE_OUT_CIRC = 1, T3D_FOG_CURVE_IN_CIRC = 2 };
enum { T3D_FILTER_NEAREST = 0, T3D_FILTER_LINEAR = 1 };
/* Light node types handled on the path (node.go:25-43) */
enum { T3D_LIGHT_AMBIENT = 0, T3D_LIGHT_POINT = 1, T3D_LIGHT_DIRECTIONAL = 2, T3D_LIGHT_VOLUME = 3 };
/* Execution forms of t3d_render */
enum {
  T3D_EXEC_PLAIN = 0,   /* one kernel chain per render pair, like one ProcessVertices+flush each */
  T3D_EXEC_BATCHED = 1  /* all render pairs of the frame in one launch chain, replayed from a HIP graph */
};

/* Matrix4 (matrix.go:9-10): m[row][col], row vectors, translation in row 3. */
typedef struct { float m[4][4]; } t3d_mat4;

/* ebiten.Vertex (ebiten v2.9.7): DstX, DstY, SrcX, SrcY, ColorR..A, Custom0..3.
 * Written by camera.go:1679-1810; fields the reference leaves untouched are 0 here. */
typedef struct {
  float dst_x, dst_y, src_x, src_y;
  float color_r, color_g, color_b, color_a;
  float custom0, custom1, custom2, custom3;
} t3d_vertex;

/* MeshPart triangle range (mesh.go:2183-2190): triangles
 * [triangle_start, triangle_start + triangle_count) of the mesh. */
typedef struct {
  uint32_t triangle_start;
  uint32_t triangle_count;
} t3d_part_range;

/* Mesh vertex SoA + triangles (mesh.go:260-302, Triangle mesh.go:1695-1711).
 * Indices are absolute mesh vertex indices (Triangle.VertexIndexA/B/C after
 * AddTriangles adds VertexIndexStart, mesh.go:2314).  Triangle.Center is
 * derived from `positions` at upload exactly as RecalculateCenter does
 * (mesh.go:1980-1982). */
typedef struct {
  uint32_t vertex_count;
  const float* positions;            /* [V][3] VertexPositions */
  const float* normals;              /* [V][3] VertexNormals */
  const float* colors;               /* [V][4] VertexColors[VertexActiveColorChannel], NULL if channel < 0 */
  const float* uvs;                  /* [V][2] VertexUVs, NULL = all zero */
  const uint16_t* bone_indices;      /* [V][4] index into the model's bone palette, NULL for unskinned meshes */
  const float* bone_weights;         /* [V][4] VertexWeights (0 = unused influence slot) */
  uint32_t shape_key_count;          /* MeshShapeKey count (mesh.go:796-822) */
  const float* shape_key_positions;  /* [K][V][3] */
  const float* shape_key_normals;    /* [K][V][3] */
  uint32_t triangle_count;
  const uint32_t* indices;           /* [T][3] */
  uint32_t part_count;
  const t3d_part_range* parts;       /* [part_count] */
  float dimensions_min[3];           /* Mesh.Dimensions (mesh.go:13-15) */
  float dimensions_max[3];
} t3d_mesh_desc;

/* Material state read by the path (material.go:66-131). */
typedef struct {
  int32_t present;                   /* MeshPart.Material != nil */
  int32_t visible;                   /* Material.Visible (mesh.go:2402-2407) */
  float color[4];                    /* Material.Color */
  int32_t backface_culling;          /* Material.BackfaceCulling */
  int32_t triangle_sort_mode;        /* T3D_SORT_* */
  int32_t shadeless;
  int32_t fogless;
  int32_t lighting_mode;             /* T3D_LIGHTING_* */
  int32_t light_volume_shading_mode; /* T3D_LV_* */
  int32_t transparency_mode;         /* T3D_TRANSPARENCY_* */
  int32_t blend_source_over;         /* Material.Blend == ebiten.BlendSourceOver */
  int32_t render_order;              /* Material.RenderOrder */
  int32_t use_texture;               /* Material.Texture != nil && Material.UseTexture */
  float texture_size[2];             /* texture Bounds().Dx(), Dy() */
  int32_t billboard;                 /* BillboardEnabled (model.go:634-693) */
  int32_t custom_depth_function;     /* CustomDepthFunction != nil: unsupported */
  int32_t depth_mode;                /* BillboardedDepthMode: T3D_DEPTH_MODE_* (unbillboarded needs a static model) */
  int32_t billboard_up;              /* BillboardUpDirection: T3D_BILLBOARD_UP_* */
  int32_t billboard_lock[3];         /* BillboardLockX, BillboardLockY, BillboardLockZ */
  /* read only by the raster sink (t3d_frame_rasterize) */
  uint64_t texture;                  /* with use_texture: handle from t3d_texture_upload (0: the default white image) */
  int32_t texture_filter;            /* Material.TextureFilterMode: T3D_FILTER_* */
  int32_t pad_texture;
} t3d_material_desc;

/* A light as seen by Camera.Render (light.go:12-30 and the concrete types). */
typedef struct t3d_light_desc {
  int32_t type;                      /* T3D_LIGHT_* */
  int32_t visible;                   /* Node IsVisible() */
  float color[3];                    /* LightBase.color RGB */
  float energy;                      /* LightBase.energy */
  float range;                       /* PointLight.Range() (light.go:269-277) */
  t3d_mat4 transform;                /* Node.Transform(): world position is row 3 */
  t3d_mat4 world_rotation;           /* Node.WorldRotation(): directional forward is row 2 */
  /* LightVolume (light.go:376-680) */
  uint64_t volume;                   /* handle from t3d_volume_upload (cell colours) */
  float volume_min[3];               /* LightVolume.dimensions */
  float volume_max[3];
  float cell_size[3];
  int32_t cell_count[3];
  float offset[3];                   /* LightVolumeOffset */
  float light_bias[3];               /* LightVolumeLightBias */
  /* LightVolume.useLightsRenderSet (light.go:392, 566-580, 1508-1513): the
   * lights its cells' useLights name, as t3d_volume_set_cell_lights indexes
   * them.  In the per-vertex shading modes each of a cell's lights runs its
   * VertexLight before the volume adds its own colour (light.go:650-654),
   * whatever its visibility and range.  Ambient, point and directional
   * lights; read only when the volume has cell light lists. */
  const struct t3d_light_desc* use_lights;
  uint32_t use_light_count;
  int32_t pad_use;
} t3d_light_desc;

/* One bone of an armature as Node state (node.go:572-608): local position,
 * scale and rotation matrix, the index of its parent bone (-1: the root
 * bones' parent node, t3d_model_desc.armature_parent), and its inverse bind
 * matrix (gltf.go:1499).  Bones are listed parents first.  The device
 * computes Transform() = Scale(scale)·rotation·Translate(position)·parent
 * Transform() and boneInfluence = inverse_bind·Transform() (node.go:592-594)
 * for every bone, each frame. */
typedef struct {
  float position[3];
  float scale[3];
  int32_t parent;
  uint32_t pad;
  t3d_mat4 rotation;
  t3d_mat4 inverse_bind;
} t3d_bone_pose;

/* AnimationTrack interpolation (animation.go:18-20).  Cubic is unimplemented
 * in the reference and samples as linear there, and here. */
enum { T3D_INTERPOLATION_LINEAR = 0, T3D_INTERPOLATION_CONSTANT = 1, T3D_INTERPOLATION_CUBIC = 2 };

/* AnimationTrack (animation.go:40-54): keyframe times in seconds, in
 * non-decreasing order, and per keyframe 3 floats (position, scale) or a
 * quaternion x, y, z, w (rotation). */
typedef struct {
  uint32_t keyframe_count;           /* 0: the channel has no such track */
  int32_t interpolation;             /* T3D_INTERPOLATION_* */
  const float* times;                /* [keyframe_count] */
  const float* values;               /* [keyframe_count] x 3 or x 4 */
} t3d_anim_track;

/* AnimationChannel (animation.go:188-200) of one bone. */
typedef struct {
  t3d_anim_track position, scale, rotation;
} t3d_anim_channel;

/* Model (model.go:20-61) with the node state the path reads. */
typedef struct {
  uint64_t mesh;                     /* handle from t3d_mesh_upload */
  t3d_mat4 transform;                /* Node.Transform() (node.go:572-608) */
  t3d_mat4 world_rotation;           /* Node.WorldRotation() (node.go:916-938) */
  float color[4];                    /* Model.Color */
  int32_t visible;
  int32_t shadeless;
  int32_t frustum_culling;           /* Model.FrustumCulling */
  int32_t auto_batch_static;         /* AutoBatchMode == AutoBatchStatic: skipped (camera.go:2069) */
  int32_t skinned;
  uint32_t bone_count;
  const t3d_mat4* bone_influences;   /* [bone_count] bone.boneInfluence = inverseBind·world (node.go:592-594) */
  const float* shape_key_weights;    /* [mesh shape_key_count] MeshShapeKey weights */
  const t3d_material_desc* materials;/* [mesh part_count], one per MeshPart; NULL = no materials */
  int32_t light_group_active;        /* LightGroup != nil && LightGroup.Active */
  uint32_t light_group_count;
  const t3d_light_desc* light_group; /* LightGroup.Lights */
  int32_t has_vertex_transform_function; /* Go callback: unsupported */
  int32_t has_vertex_clip_function;      /* Go callback: unsupported */
  /* Skinned models may give their bones' poses instead of bone_influences:
   * [bone_count] poses, parents first; the palette is then computed on the
   * device each frame (bone_influences is ignored). */
  const t3d_bone_pose* bone_poses;
  const t3d_mat4* armature_parent;   /* Transform() of the root bones' parent node; NULL: roots have no parent */
  /* With bone_poses: an animation from t3d_animation_upload (0: none) whose
   * channel b animates bone b, sampled at animation_time (the player's
   * playhead when AnimationPlayer.Update samples, animation.go:674-720):
   * each present track replaces the bone's position, scale or rotation
   * (animation.go:902-983, no blend, absolute motion). */
  uint64_t animation;
  float animation_time;
  uint32_t pad_animation;
  /* Dynamic batching (Model.DynamicBatchAdd, model.go:217-247).  With
   * dynamic_batched set, this model is in the dynamic batch of
   * models[batch_owner] under that model's MeshPart batch_part
   * (Model.DynamicBatchOwner != nil): it is no render pair of its own and
   * draws only inside its owner's flush.  batch_order is its position in the
   * owner's DynamicBatchModels[part] slice (the order DynamicBatchAdd appended
   * it).  A model named as owner is a dynamic batcher: its own mesh is not
   * drawn, and each of its parts with batched models is one render pair
   * (camera.go:1402-1453) whose flush holds every part of its visible batched
   * models, farthest from the camera first (camera.go:2075-2099).  Parts of
   * one owner go in ascending part order (the reference walks a Go map).  A
   * batched model's own auto_batch_static and hidden materials are not
   * consulted, as in the reference. */
  int32_t dynamic_batched;
  uint32_t batch_owner;
  uint32_t batch_part;
  uint32_t batch_order;
} t3d_model_desc;

/* Camera state (camera.go:256-321, 324-547). */
typedef struct {
  int32_t width, height;             /* result texture size */
  int32_t perspective;
  float fov;                         /* vertical FOV in degrees */
  float near, far;
  float ortho_scale;
  t3d_mat4 transform;                /* Node.Transform(): world position is row 3 */
  t3d_mat4 world_rotation;           /* Node.WorldRotation() at the last Clear() (camera.go:1085-1088) */
  int32_t render_depth;              /* Camera.RenderDepth */
  int32_t render_normals;            /* Camera.RenderNormals: also fill the normal vertex list */
  int32_t perspective_corrected_texture_mapping;
  float depth_margin;                /* Camera.DepthMargin */
  float vertex_snapping;             /* Camera.VertexSnapping */
  int32_t max_light_count;           /* Camera.MaxLightCount */
  int32_t has_max_light_count_focus;
  float max_light_count_focus[3];    /* MaxLightCountFocusPoint world position */
} t3d_camera_desc;

/* World (world.go:25-42). present = 0 means scene.World == nil. */
typedef struct {
  int32_t present;
  int32_t lighting_on;
  int32_t fog_on;
  int32_t fog_mode;                  /* T3D_FOG_* */
  float fog_color[4];
  float fog_range[2];
  float ambient_color[3];            /* World.AmbientLight */
  float ambient_energy;
  int32_t ambient_visible;
  /* read only by the raster sink */
  float clear_color[4];              /* World.ClearColor, what Camera.Clear fills the colour texture with (camera.go:1002-1013) */
  int32_t fog_curve;                 /* World.FogCurve: T3D_FOG_CURVE_* */
  float dithered_fog_size;           /* World.DitheredFogSize (0: no dithering) */
} t3d_world_desc;

/* One Camera.Render(scene, lights, models) call. */
typedef struct {
  const t3d_camera_desc* camera;
  const t3d_world_desc* world;
  uint32_t light_count;
  const t3d_light_desc* lights;      /* the `lights` NodeIterator, in order */
  uint32_t model_count;
  const t3d_model_desc* models;      /* the `models` NodeIterator, in order */
  int32_t exec_mode;                 /* T3D_EXEC_* */
  /* 1: every render pair runs ProcessVertices with processOnlyVisible false
   * (model.go:600-960), as BakeLighting and the debug views do (model.go:1233,
   * camera.go:2434, 2528, 2581): no alpha-0 early out, and every triangle is
   * kept (no near/far, screen or backface test, NaN and infinite depth keys
   * kept; a NaN key, which the reference's bucket index int(NaN) panics on,
   * sorts into bin 0).  0: Camera.Render's processing. */
  int32_t process_all_triangles;
} t3d_render_desc;

/* One render pair = one ProcessVertices + flush (camera.go:2057-2109).  For a
 * dynamic batch the pair is the owner's (model_index = owner, part_index =
 * its batched part), and the flush's vertex lists hold the batched models'
 * parts one after another, indices continuing across them as the
 * reference's vertexListIndex does; triangle_count then sums their parts. */
typedef struct {
  uint32_t model_index;              /* index into t3d_render_desc.models */
  uint32_t part_index;               /* MeshPart index within the mesh */
  uint32_t triangle_count;           /* MeshPart.TriangleCount() */
  uint32_t survivor_count;           /* triangles drawn: the flush's vertexListIndex / 3 */
  uint64_t vertex_offset;            /* first vertex of the pair in the frame's vertex arrays */
} t3d_pair_result;

typedef struct t3d_ctx t3d_ctx;
typedef struct t3d_frame t3d_frame;

/* ---- context --------------------------------------------------------- */

/* Create a context on HIP device `device`.  hbm_budget_bytes bounds the
 * device memory the context may hold (meshes + frames); 0 = no bound. */
t3d_status t3d_ctx_create(int32_t device, uint64_t hbm_budget_bytes, t3d_ctx** out);
t3d_status t3d_ctx_destroy(t3d_ctx* ctx);
const char* t3d_last_error(const t3d_ctx* ctx);
int32_t t3d_abi_version(void);
/* Device bytes currently held by the context. */
uint64_t t3d_ctx_memory_used(const t3d_ctx* ctx);

/* ---- resources (resident in HBM) ------------------------------------- */

/* Upload a Mesh; replaces the reference's per-frame reads of Mesh.Vertex*
 * slices (mesh.go:274-290). */
t3d_status t3d_mesh_upload(t3d_ctx* ctx, const t3d_mesh_desc* desc, uint64_t* out_handle);
/* Free a mesh.  T3D_ERR_BUSY while a submitted job still references it. */
t3d_status t3d_mesh_free(t3d_ctx* ctx, uint64_t handle);
/* Upload LightVolume cell colours: [cell_count] RGBA float32, cell order
 * x + y*cx + z*cx*cy (getLightCell, light.go:1267-1272). */
t3d_status t3d_volume_upload(t3d_ctx* ctx, const float* cells_rgba, uint32_t cell_count, uint64_t* out_handle);
/* T3D_ERR_BUSY while a submitted job still references the volume. */
t3d_status t3d_volume_free(t3d_ctx* ctx, uint64_t handle);
/* Every cell's useLights list at once (LightVolumeCell.SetUseLights and
 * ClearUseLights, light.go:1508-1518): cell i names the lights
 * light_index[cell_first[i] .. cell_first[i+1]), in order and repeats kept,
 * each an index into the volume light's t3d_light_desc.use_lights.
 * cell_first holds cell_count + 1 non-decreasing entries from 0; a null
 * cell_first clears every list.  T3D_ERR_BUSY while a submitted job still
 * references the volume. */
t3d_status t3d_volume_set_cell_lights(t3d_ctx* ctx, uint64_t handle, const uint32_t* cell_first,
                                      const uint32_t* light_index);

/* An Animation's channels (animation.go:188-260), resident on the device;
 * replaces the per-frame AnimationTrack.ValueAsVector / ValueAsQuaternion
 * calls of AnimationPlayer.updateValues (animation.go:63-160, 674-720).
 * Times must be finite and non-decreasing within a track. */
t3d_status t3d_animation_upload(t3d_ctx* ctx, const t3d_anim_channel* channels, uint32_t channel_count,
                                uint64_t* out_handle);
/* T3D_ERR_BUSY while a submitted job still references the animation. */
t3d_status t3d_animation_free(t3d_ctx* ctx, uint64_t handle);

/* ---- frames ---------------------------------------------------------- */

/* A frame owns the device output arrays of one Camera.Render call and the
 * per-frame scratch.  It is reused across renders; outputs stay valid until
 * the next t3d_render into the same frame. */
t3d_status t3d_frame_create(t3d_ctx* ctx, t3d_frame** out);
t3d_status t3d_frame_destroy(t3d_ctx* ctx, t3d_frame* frame);

/* Camera.Render (camera.go:1269): runs the whole geometry path on the device
 * and waits for it.  On return the frame holds, per render pair in the
 * reference's draw order (solids in traversal order, then transparents
 * sorted by RenderOrder and distance, camera.go:2050-2109), the colour
 * vertices, depth vertices and indices that flush() would draw. */
t3d_status t3d_render(t3d_ctx* ctx, t3d_frame* frame, const t3d_render_desc* desc);

/* The same render without the wait: builds the frame's tables, enqueues the
 * table upload (on the context's copy stream) and the kernel chain, and
 * returns.  A caller that alternates two frames overlaps the host work of
 * one Camera.Render with the device work of the previous one.  Rendering
 * into a frame that is still in flight waits for it first; its earlier
 * results are then replaced. */
t3d_status t3d_render_async(t3d_ctx* ctx, t3d_frame* frame, const t3d_render_desc* desc);
/* Waits for the frame's last t3d_render_async and publishes its results
 * (pairs, survivor counts).  T3D_OK at once when nothing is in flight. */
t3d_status t3d_frame_wait(t3d_ctx* ctx, t3d_frame* frame);

/* Number of render pairs of the last render and their results
 * (T3D_ERR_NOT_READY while an asynchronous render is in flight). */
t3d_status t3d_frame_pairs(const t3d_frame* frame, uint32_t* out_count, const t3d_pair_result** out_pairs);
/* Device pointers of the frame's output arrays (valid until the next
 * render into this frame).  Pair p occupies vertices
 * [vertex_offset, vertex_offset + 3*survivor_count). */
t3d_status t3d_frame_device_outputs(const t3d_frame* frame, const t3d_vertex** color_vertices,
                                    const t3d_vertex** depth_vertices, const uint32_t** indices);
/* Copy all pairs' outputs, concatenated in pair order, to host memory.
 * Buffers must hold 3*sum(survivor_count) entries each; any may be NULL.
 * These accessors return T3D_ERR_NOT_READY while an asynchronous render of
 * the frame is in flight (t3d_frame_wait first). */
t3d_status t3d_frame_download(t3d_ctx* ctx, const t3d_frame* frame, t3d_vertex* color_vertices,
                              t3d_vertex* depth_vertices, uint32_t* indices);
/* Camera.RenderNormals: the normalVertexList each flush draws into the
 * normal texture (camera.go:1710-1722, 1989-1992), laid out like the colour
 * vertices (device pointer; T3D_ERR_NOT_FOUND when the last render's camera
 * does not render normals), and its copy to host memory, concatenated in
 * pair order like t3d_frame_download. */
t3d_status t3d_frame_device_normals(const t3d_frame* frame, const t3d_vertex** normal_vertices);
t3d_status t3d_frame_download_normals(t3d_ctx* ctx, const t3d_frame* frame, t3d_vertex* normal_vertices);
/* ---- raster sink ------------------------------------------------------
 * The flushes of a rendered frame drawn into the camera's result textures on
 * the GPU, as flush hands them to ebiten (camera.go:1829-2048) with the
 * depth, clip-alpha and base3d shaders (camera.go:350-531,
 * shaders/base3d.kage:61-153): Camera.Clear with the world's clear colour,
 * then per flush the depth pass into a cleared intermediate tested against
 * the result depth, the intermediate drawn over the result depth unless the
 * pair is transparent, and the colour pass (texture, vertex colour, fog,
 * dithering); without RenderDepth one textured draw per flush.  Images are
 * W x H RGBA8 with premultiplied alpha, rows top first.  The fixed-function
 * rules ebiten's GPU applies (pixel centres, top-left ties, linear
 * interpolation, 8-bit rounding) are fixed by this library.  Frames of one
 * Camera.Render only; flushes blending other than SourceOver are refused.
 * (The reference's normal texture is not drawn: its normal vertices never
 * get an alpha, camera.go:1710-1722, so that draw leaves it cleared.) */
/* An ebiten.Image for Material.Texture: RGBA8, premultiplied, row-major.
 * A texture must stay allocated while frames that name it are rasterized. */
t3d_status t3d_texture_upload(t3d_ctx* ctx, uint32_t width, uint32_t height, const uint8_t* rgba8, uint64_t* out_handle);
t3d_status t3d_texture_free(t3d_ctx* ctx, uint64_t handle);
/* Draw the last render of `frame` (finished: after t3d_render, or after
 * t3d_frame_wait) into its images; returns once they are complete in HBM. */
t3d_status t3d_frame_rasterize(t3d_ctx* ctx, t3d_frame* frame);
/* The images of the last t3d_frame_rasterize: device pointers (depth NULL
 * without RenderDepth), valid until the frame's next rasterize; and their
 * copy to host memory (W * H * 4 bytes each; depth may be NULL). */
t3d_status t3d_frame_device_images(const t3d_frame* frame, const uint8_t** color, const uint8_t** depth);
t3d_status t3d_frame_download_images(t3d_ctx* ctx, const t3d_frame* frame, uint8_t* color, uint8_t* depth);
/* Mesh triangle index of every drawn triangle, pair-concatenated in draw
 * order: the order sortingTriangleBucket.ForEach visits them
 * (sortingTriangles.go:93-125).  Holds sum(survivor_count) entries. */
t3d_status t3d_frame_download_triangles(t3d_ctx* ctx, const t3d_frame* frame, uint32_t* triangle_ids);
/* Wall-clock milliseconds of the last t3d_render (host prologue + device
 * work + survivor readback). */
double t3d_frame_last_ms(const t3d_frame* frame);
/* Duration in milliseconds of the last render's device work, measured with
 * HIP events on the context's stream; per_kernel receives up to
 * max_kernels (name, total ms, launches) rows. */
typedef struct {
  char name[32];
  double total_ms;
  uint32_t launches;
  uint64_t algorithmic_bytes;        /* algorithmic HBM bytes the launches of this kernel moved */
} t3d_kernel_time;
t3d_status t3d_profile_enable(t3d_ctx* ctx, int32_t on);
t3d_status t3d_profile_read(t3d_ctx* ctx, t3d_kernel_time* per_kernel, uint32_t max_kernels,
                            uint32_t* out_count, int32_t reset);

/* ---- rendering without allocation ----------------------------------- */

/* Bytes of the static device scratch t3d_render_noalloc carves a frame's
 * tables and scratch from (one array per GPU, held until the process ends). */
#define T3D_NOALLOC_SCRATCH_BYTES (64ull << 20)

/* No counterpart in the reference.  For callers that cannot allocate device
 * memory (tight serving loops, harnesses that forbid hipMalloc): the same
 * Camera.Render as t3d_render (same kernels, same results), but the frame's
 * tables and scratch come from a static per-GPU device array of
 * T3D_NOALLOC_SCRATCH_BYTES and the outputs are the caller's device arrays,
 * each holding vertex_capacity entries.  Calls no hipMalloc, hipHostMalloc or
 * hipFree and builds no graph; the tables go up from pageable host memory
 * and the call waits for the frame.  Pair p's results occupy vertices
 * [vertex_offset, vertex_offset + 3*survivor_count) of the caller's arrays
 * (t3d_frame_pairs), so the arrays must hold 3 x the triangles of every drawn
 * render pair (T3D_ERR_INVALID_ARGUMENT otherwise, naming the count); a frame
 * whose tables and scratch exceed the array is refused with T3D_ERR_BUDGET.
 * The frame argument carries only host bookkeeping here.  Every call on one
 * GPU shares its array, from any context: calls that overlap on one GPU race. */
t3d_status t3d_render_noalloc(t3d_ctx* ctx, t3d_frame* frame, const t3d_render_desc* desc,
                              t3d_vertex* color_vertices, t3d_vertex* depth_vertices, uint32_t* indices,
                              uint64_t vertex_capacity);

/* ---- tuning ---------------------------------------------------------- */

/* Launch shape of the batched form: chosen by t3d_autotune or set by hand. */
typedef struct {
  uint32_t vertex_block;             /* threads per block, vertex kernel */
  uint32_t vertices_per_thread;
  uint32_t triangle_tile;            /* triangles per tile in the cull/sort kernels (64..2048) */
  uint32_t scatter_waves;            /* wavefronts per tile in the sort's scatter pass: 1 or 4 */
  uint32_t use_graph;                /* replay the batched chain from a HIP graph */
  uint32_t store_mode;               /* assembly stores: bit 0 = direct per-lane stores instead of
                                        LDS-staged rows, bit 1 = non-temporal (streaming) stores,
                                        bit 2 = stage 32 rows at a time instead of 64, bit 3 =
                                        software-pipelined gathers, bit 4 = four wavefronts per block,
                                        bit 5 = one output stream per block (colour vertices and indices,
                                        or depth vertices; twice the blocks), bit 6 = the assembly's
                                        blocks in launch order whatever xcd_swizzle says for the other
                                        passes (0-4, 6, 8, 10, 12, 14, 16, 18, 20, 22, 34, 38, 50, 54,
                                        and 80, 82, 114 for 16, 18, 50 with bit 6) */
  uint32_t assemble_split;           /* assembly blocks per tile: a power of two, tile/split >= 64 */
  uint32_t xcd_swizzle;              /* 1: each XCD takes a contiguous range of a launch's blocks */
  uint32_t vertex_split;             /* 1: clip positions and colours in two arrays instead of one
                                        interleaved 32-byte record per vertex */
  uint32_t cull_block;               /* threads per tile block in the cull pass: 256, 512 or 1024 */
  uint32_t model_stage;              /* 1: the batched form runs the model-level stage (frustum test, render
                                        pairs and their order, per-pair slices) on the device when the
                                        frame allows it: no dynamic batch, RenderDepth on, every model at a
                                        finite distance from its camera; 0: always on the host */
} t3d_tuning;
t3d_status t3d_get_tuning(const t3d_ctx* ctx, t3d_tuning* out);
t3d_status t3d_set_tuning(t3d_ctx* ctx, const t3d_tuning* tuning);
/* Time candidate launch shapes on `desc` (coordinate descent, two sweeps).
 * Each shape is timed as the batched form runs, with two frames in flight
 * (the given frame and a second one the call creates and destroys), and by
 * its synchronous frame; medians of `reps` samples; synchronous frames only
 * when the second frame does not fit.  A change is kept when it is at least
 * 1 % better on one time and no worse on the other (pipelined within 0.5 %,
 * or 1 % for a synchronous frame 10 % shorter; synchronous within 2 %).
 * Writes the winner to *out (may be NULL).  The given frame holds the last
 * candidate's results afterwards. */
t3d_status t3d_autotune(t3d_ctx* ctx, t3d_frame* frame, const t3d_render_desc* desc, uint32_t reps,
                        t3d_tuning* out);

/* ---- glTF -> resident mesh pool ------------------------------------------
 * LoadGLTFData's mesh and skin reading (gltf.go:560-840, 1449-1505): a glTF 2.0
 * document (GLB, or JSON with base64 data URIs) decoded on the host into the
 * t3d_mesh_desc arrays the reference's Mesh holds after loading (one MeshPart
 * per primitive, vertices appended, indices offset by the part's first vertex,
 * UV V flipped, COLOR_0 as 8-bit sRGB-converted colours that are the active
 * channel when the mesh extras name its colour channels, influences of weight
 * <= 0 dropped, morph targets as shape keys, dimensions over the positions),
 * and the skins' joints and inverse bind matrices.  t3d_gltf_upload puts every
 * mesh in the context's resident pool.  Refused (T3D_ERR_INVALID_ARGUMENT,
 * message in err): sparse accessors, external buffer files, primitives without
 * indices or with an index count that is not a multiple of three, morph
 * targets on a mesh of several primitives, auto-subdivided meshes. */
typedef struct t3d_gltf t3d_gltf;
typedef struct {
  char name[64];                     /* glTF mesh name (truncated) */
  uint32_t vertex_count, triangle_count, part_count, shape_key_count;
  int32_t has_colors;                /* VertexActiveColorChannel is 0 (else -1: desc.colors is NULL) */
  int32_t skinned;                   /* JOINTS_0 / WEIGHTS_0 present */
} t3d_gltf_mesh_info;
/* Needs no context and no GPU.  err (may be NULL) receives the reason of a refusal. */
t3d_status t3d_gltf_parse(const void* data, uint64_t size, t3d_gltf** out, char* err, uint32_t err_len);
void t3d_gltf_free(t3d_gltf* gltf);
uint32_t t3d_gltf_mesh_count(const t3d_gltf* gltf);
/* Mesh i: its summary, its arrays as t3d_mesh_upload takes them (pointers into
 * the document, valid until t3d_gltf_free), and per part the glTF material
 * index of its primitive (-1: none).  Any output may be NULL. */
t3d_status t3d_gltf_mesh(const t3d_gltf* gltf, uint32_t i, t3d_gltf_mesh_info* info, t3d_mesh_desc* desc,
                         const int32_t** part_materials);
uint32_t t3d_gltf_skin_count(const t3d_gltf* gltf);
/* Skin i: its joint node indices (bone b of a skinned mesh's JOINTS_0 is joint
 * b) and their inverse bind matrices as Matrix4 (column c = the c-th four
 * floats of the accessor element, gltf.go:1487-1497). */
t3d_status t3d_gltf_skin(const t3d_gltf* gltf, uint32_t i, uint32_t* joint_count, const int32_t** joints,
                         const t3d_mat4** inverse_bind);
/* Every mesh of the document into the resident pool: handles[i] for mesh i
 * (handles holds t3d_gltf_mesh_count entries).  On failure none stays resident. */
t3d_status t3d_gltf_upload(t3d_ctx* ctx, const t3d_gltf* gltf, uint64_t* handles);

/* ---- serving ----------------------------------------------------------- */

/* A render-request loop over a context's resident meshes.  Jobs (one
 * Camera.Render each: camera + a subset of models + lights) are admitted by
 * an HBM budget and batched continuously: whenever one of `slots` launch
 * slots is free, every admitted job waiting at that moment (up to
 * max_batch_jobs) goes into one launch chain on that slot's HIP stream, so
 * batches overlap on the device while new requests queue.  Every job's
 * results are identical to rendering it alone with t3d_render.  All
 * functions below are thread-safe. */
typedef struct t3d_server t3d_server;
typedef struct {
  uint32_t slots;                    /* batches in flight at once, each on its own stream (>= 1) */
  uint32_t max_batch_jobs;           /* jobs merged into one launch chain (>= 1) */
  uint64_t hbm_budget_bytes;         /* frame memory in-flight jobs may hold; 0 = unbounded */
  int32_t auto_release;              /* release a job's outputs as soon as it completes */
} t3d_server_config;
enum { T3D_JOB_QUEUED = 0, T3D_JOB_RUNNING = 1, T3D_JOB_DONE = 2, T3D_JOB_FAILED = 3, T3D_JOB_CANCELLED = 4 };
typedef struct {
  int32_t state;                     /* T3D_JOB_* */
  t3d_status status;                 /* error of a failed job */
  double submit_ms, start_ms, done_ms; /* steady-clock times since server creation */
  uint32_t pair_count;
  uint64_t triangles;                /* input triangles of the job's models */
  uint64_t survivors;                /* triangles drawn */
  uint64_t batch;                    /* batch the job ran in */
  uint32_t batch_jobs;               /* jobs in that batch */
  uint64_t admitted_bytes;           /* frame memory the job counts against the budget */
} t3d_job_info;
typedef struct {
  uint64_t submitted, completed, failed, rejected, cancelled, batches;
  uint64_t triangles;                /* input triangles of completed jobs */
  uint32_t queue_depth, running_batches;
  uint64_t memory_in_flight;         /* admitted bytes of queued-for-launch and running jobs */
  double p50_ms, p99_ms;             /* submit -> outputs in HBM, completed jobs */
  double first_submit_ms, last_done_ms;
} t3d_server_stats;

t3d_status t3d_server_create(t3d_ctx* ctx, const t3d_server_config* cfg, t3d_server** out);
/* Waits for running batches, cancels queued jobs, frees everything. */
t3d_status t3d_server_destroy(t3d_server* s);
/* Copies the descriptor (and everything it points to); the caller may reuse
 * its memory on return.  T3D_ERR_BUDGET if the job alone exceeds the budget. */
t3d_status t3d_job_submit(t3d_server* s, const t3d_render_desc* desc, uint64_t* out_job);
/* T3D_OK when done, T3D_ERR_NOT_READY while queued or running, the job's
 * error when it failed, T3D_ERR_CANCELLED when cancelled. */
t3d_status t3d_job_poll(t3d_server* s, uint64_t job);
t3d_status t3d_job_wait(t3d_server* s, uint64_t job, uint32_t timeout_ms);
t3d_status t3d_job_info_get(t3d_server* s, uint64_t job, t3d_job_info* out);
/* Results of a completed job, as t3d_frame_pairs + t3d_frame_download(_triangles)
 * give them for t3d_render (vertex offsets count from the job's first vertex).
 * pairs holds pair_count entries; the arrays 3*survivors (tri_ids: survivors). */
t3d_status t3d_job_download(t3d_server* s, uint64_t job, t3d_pair_result* pairs, t3d_vertex* color_vertices,
                            t3d_vertex* depth_vertices, uint32_t* indices, uint32_t* triangle_ids);
/* The normal vertex list of a completed job whose camera renders normals
 * (3*survivors entries; T3D_ERR_NOT_FOUND otherwise). */
t3d_status t3d_job_download_normals(t3d_server* s, uint64_t job, t3d_vertex* normal_vertices);
/* Frees a finished job's outputs (its batch's memory is reused once all of
 * the batch's jobs are released). */
t3d_status t3d_job_release(t3d_server* s, uint64_t job);
/* Cancels a queued job; T3D_ERR_BUSY once it is running or finished. */
t3d_status t3d_job_cancel(t3d_server* s, uint64_t job);
/* Waits until no job is queued or running. */
t3d_status t3d_server_drain(t3d_server* s, uint32_t timeout_ms);
t3d_status t3d_server_stats_get(t3d_server* s, t3d_server_stats* out);
/* Forget latency samples and counters (e.g. after warm-up). */
t3d_status t3d_server_stats_reset(t3d_server* s);

/* ---- self checks ------------------------------------------------------ */

/* word != 0: before every render of this context, fill the frame's output
 * arrays with `word`, so a vertex or index the path fails to write shows up
 * instead of keeping an earlier render's value (tests pass a NaN pattern).
 * 0 turns it off.  No counterpart in the reference: a debugging aid. */
t3d_status t3d_debug_poison_outputs(t3d_ctx* ctx, uint32_t word);

/* Bounding spheres of n models as the frustum test computes them
 * (Model.onTransformUpdate, model.go:142-182): out_batched from the
 * four-models-at-a-time path t3d_render uses, out_scalar from the one-model
 * path, 4 floats per model (centre x, y, z, radius).  dims holds 6 floats per
 * model (mesh dimension min xyz, max xyz), skinned one byte per model.  Needs
 * no context or device.  Debugging aid with no reference counterpart: lets a
 * host-only test check that the two paths agree bit for bit. */
t3d_status t3d_debug_model_spheres(const t3d_mat4* transforms, const float* dims, const uint8_t* skinned, uint32_t n,
                                   float* out_batched, float* out_scalar);

/* Runs the device's float32 primitives on n inputs: out[0..n) = x/y,
 * out[n..2n) = sqrt(|x|), out[2n..3n) = x*y + x (unfused), out[3n..4n) =
 * round-half-away(x).  Lets a test confirm the device rounds exactly as the
 * host does.  Host pointers; n <= 1<<20. */
t3d_status t3d_selftest_fp(t3d_ctx* ctx, const float* x, const float* y, float* out, uint32_t n);
/* Copies `bytes` (a multiple of 16) between two device buffers `reps` times
 * with a plain float4 kernel and returns the mean kernel time in *out_ms:
 * the reference point that HBM byte counters are scaled against. */
t3d_status t3d_selftest_copy(t3d_ctx* ctx, uint64_t bytes, uint32_t reps, double* out_ms);
/* sizeof of each public struct, for binding checks: fills up to max entries
 * in the order vertex, mesh_desc, material_desc, light_desc, model_desc,
 * camera_desc, world_desc, render_desc, pair_result, kernel_time, tuning,
 * server_config, job_info, server_stats, bone_pose, anim_track, anim_channel. */
uint32_t t3d_struct_sizes(uint32_t* sizes, uint32_t max);

#ifdef __cplusplus
}
#endif

#endif /* T3D_H */
