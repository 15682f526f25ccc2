/*
 * t3d_oracle.h — CPU restatement of Tetra3D's per-frame geometry path.
 *
 * TEST INFRASTRUCTURE ONLY.  This is the checker for the HIP path: tests,
 * the smoke check and the CPU-baseline leg of the benchmark load it; the
 * product library never links or calls it.
 *
 * It restates, single-threaded and in float32 with the reference's operation
 * order (built with -ffp-contract=off), the CPU code that Camera.Render runs
 * (camera.go:1269-1827, model.go:142-182, 559-995, sortingTriangles.go,
 * light.go, matrix.go, vector.go, math32.go).  Scratch buffers are sized to
 * the mesh, where the reference grows fixed-size globals (tetra3d.go:41-90);
 * that changes no arithmetic.
 */
#ifndef T3D_ORACLE_H
#define T3D_ORACLE_H

#include "../include/t3d.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct t3do_ctx t3do_ctx;

typedef struct {
  uint32_t pair_count;
  t3d_pair_result* pairs;        /* vertex_offset indexes the concatenated arrays */
  uint64_t vertex_count;         /* 3 * total survivors */
  t3d_vertex* color;             /* colorVertexList per flush, concatenated */
  t3d_vertex* depth;             /* depthVertexList per flush, concatenated */
  uint32_t* indices;             /* indexList per flush (pair-local), concatenated */
  uint32_t* triangle_ids;        /* mesh triangle index of each drawn triangle, draw order */
  float* depth_keys;             /* bucket depth of each drawn triangle */
  t3d_vertex* normals;           /* normalVertexList per flush when Camera.RenderNormals, else NULL */
} t3do_result;

t3do_ctx* t3do_create(void);
void t3do_destroy(t3do_ctx* ctx);
/* Register a mesh / light volume under a caller-chosen handle (copies data). */
int t3do_mesh_register(t3do_ctx* ctx, uint64_t handle, const t3d_mesh_desc* desc);
int t3do_volume_register(t3do_ctx* ctx, uint64_t handle, const float* cells_rgba, uint32_t cell_count);
/* Every cell's useLights (LightVolumeCell.SetUseLights, light.go:1508-1513) as t3d_volume_set_cell_lights
 * takes them; a NULL cell_first clears them. */
int t3do_volume_set_cell_lights(t3do_ctx* ctx, uint64_t handle, const uint32_t* cell_first, const uint32_t* light_index);
int t3do_animation_register(t3do_ctx* ctx, uint64_t handle, const t3d_anim_channel* channels, uint32_t channel_count);
/* Camera.Render restated.  Fills *out (free with t3do_result_free). */
int t3do_render(t3do_ctx* ctx, const t3d_render_desc* desc, t3do_result* out);
void t3do_result_free(t3do_result* out);
const char* t3do_last_error(const t3do_ctx* ctx);
/* Raster sink restated: a texture under a caller-chosen handle (RGBA8, premultiplied), and the flushes of a
 * render drawn into W*H*4-byte colour and (RenderDepth) depth images. */
int t3do_texture_register(t3do_ctx* ctx, uint64_t handle, uint32_t width, uint32_t height, const uint8_t* rgba8);
int t3do_rasterize(t3do_ctx* ctx, const t3d_render_desc* desc, const t3do_result* res, uint8_t* color, uint8_t* depth);

/* ---- unit-level entry points (used by the tests that pin the oracle) ---- */
void t3do_mat4_mult(const t3d_mat4* a, const t3d_mat4* b, t3d_mat4* out);
/* boneInfluence of every bone of an armature from its poses (node.go:572-608);
 * root may be NULL.  Returns 0, or -1 for a parent that is not an earlier bone. */
int t3do_bone_palettes(const t3d_bone_pose* poses, uint32_t n, const t3d_mat4* root, t3d_mat4* out);
/* AnimationPlayer sampling at `time` applied to poses in place (animation.go:63-160, 902-983). */
void t3do_animate_poses(const t3d_anim_channel* ch, uint32_t nch, float time, t3d_bone_pose* poses, uint32_t n);
/* Go's float64 math.Sin and math.Atan2 as restated; Quaternion.Slerp and ToMatrix4. */
double t3do_go_sin(double x);
double t3do_go_atan2(double y, double x);
void t3do_quat_slerp(const float a[4], const float b[4], float percent, float out[4]);
void t3do_quat_to_mat4(const float q[4], t3d_mat4* out);
void t3do_mat4_inverted(const t3d_mat4* a, t3d_mat4* out);
void t3do_mat4_transposed(const t3d_mat4* a, t3d_mat4* out);
void t3do_mat4_decompose(const t3d_mat4* a, float pos[3], float scale[3], t3d_mat4* rot);
void t3do_mat4_mult_vec(const t3d_mat4* a, const float v[3], float out[3]);
void t3do_mat4_mult_vec_w(const t3d_mat4* a, const float v[3], float out[4]);
void t3do_projection_perspective(float fovy, float near, float far, float w, float h, t3d_mat4* out);
void t3do_view_matrix(const t3d_camera_desc* cam, t3d_mat4* out);
void t3do_vp_matrix(const t3d_camera_desc* cam, t3d_mat4* out);
void t3do_vec3_unit(const float v[3], float out[3]);
int t3do_sphere_in_frustum(const t3d_camera_desc* cam, const float pos[3], float radius);
void t3do_model_frustum_sphere(const t3d_mat4* transform, int skinned, const float dims_min[3],
                               const float dims_max[3], float pos[3], float* radius);
void t3do_inverted_cam_pos(const t3d_mat4* model_transform, const float cam_pos[3], float out[3]);
/* sortingTriangleBucket.Sort + ForEach on `n` (depth) entries: writes the
 * entry order ForEach visits into order[n]; returns n. */
uint32_t t3do_bucket_sort(const float* depths, uint32_t n, float min_depth, float max_depth, int32_t sort_mode,
                          uint32_t* order);
uint32_t t3do_bucket_bin(float depth, float min_depth, float max_depth, int32_t sort_mode);
/* Go's sort.SliceStable on an int array with a comparator given as a
 * precomputed n*n less-table (less[i*n+j] = less(i, j) on original items);
 * permutes items[] in place.  For checking the restated algorithm. */
void t3do_go_stable_sort_table(int32_t* items, uint32_t n, const uint8_t* less_table);

#ifdef __cplusplus
}
#endif

#endif
