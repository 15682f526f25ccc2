// t3d_internal.h — device-side layout of one frame of the geometry path.
//
// HBM layout (all arrays 256-byte aligned):
//   mesh pool (resident, per mesh): pos[V]xyz f32, nrm[V]xyz f32, col[V]rgba f32,
//     uv[V] f32x2, bones[V] u16x4, weights[V] f32x4, shape keys [K][V]xyz,
//     idx[T] u32x3 — attribute-level SoA, so every per-vertex stream is a
//     unit-stride coalesced load.
//   frame tables (uploaded per render): PartGPU[P], LightGPU[L], PartLight[sum],
//     block->part prefix arrays.
//   frame scratch: per vertex a clip f32x4 and a colour f32x4, interleaved
//     (one 32-byte record per vertex) or as two arrays (the cull pass then
//     reads only clip positions), vpos[sumV] f32x4 (skinned positions),
//     key[sumT] f32,
//     tri_rec/tri_vo[sumT] u32x4 (survivor records by triangle / by slot),
//     stats[P] u32x4, hist, bsum.
//   outputs: color[3*sumT] t3d_vertex, depth[3*sumT] t3d_vertex, idx[3*sumT] u32;
//     render pair p owns vertices [out_base, out_base + 3*T_p) and fills the
//     first 3*S_p of them.
#pragma once

#include <stdint.h>

#include "t3d_math.h"

namespace t3d {

constexpr uint32_t kBins = 512;        // sortingTriangles.go:166-168
constexpr uint32_t kMaxLights = 64;    // lights of one pair the vertex pass stages in LDS (more are read from HBM)
constexpr uint32_t kMaxLdsBones = 64;  // bone palettes staged in LDS up to this size
constexpr uint32_t kScanBlock = 2048;  // elements per scan block (256 threads x 8)
constexpr uint32_t kMaxTile = 2048;    // triangles per tile (keys of a tile are staged in LDS)
constexpr uint32_t kScanPartMax = 65536;  // histogram matrices up to this size are scanned by one block

// PartGPU.flags
enum : uint32_t {
  PF_SKINNED = 1u << 0,
  PF_BACKFACE = 1u << 1,
  PF_LIGHTING = 1u << 2,
  PF_HAS_COLOR = 1u << 3,
  PF_HAS_UV = 1u << 4,
  PF_SHAPE_KEYS = 1u << 5,
  PF_NEED_RD = 1u << 6,  // a directional light lights this static part: k_uniforms inverts its rotation
  PF_UNBB = 1u << 7,     // DepthModeUnbillboarded: the depth list takes Z of Model.Transform()·VP, unsnapped
  PF_SORT_SHIFT = 8,     // 2 bits: T3D_SORT_*
  PF_LMODE_SHIFT = 10,   // 2 bits: T3D_LIGHTING_*
  PF_LVMODE_SHIFT = 12,  // 2 bits: T3D_LV_*
  PF_ALL = 1u << 14,     // ProcessVertices with processOnlyVisible false: every triangle kept (model.go:883-960)
};

// Per render pair: uniforms (camera.go:1507-1601, model.go:600-695) and the
// pair's slices of the frame arrays.
struct alignas(16) PartGPU {
  M4 mvp;              // uploaded as the base transform (Model.Transform(), or its billboarded form, model.go:634-693);
                       // k_uniforms turns it into base·VP (model.go:695)
  M4 rd;               // DirectionalLight working rotation: Transposed(Inverted(WorldRotation)) (light.go:326-330);
                       // uploaded as the model's WorldRotation, k_uniforms turns it into this
  M4 model;            // Model.Transform() (LightVolume lookups, light.go:610-628)
  float mp[4];         // mpColor = Model.Color ⊙ Material.Color (camera.go:1597-1601)
  float icp[3];        // invertedCamPos (model.go:629-630), computed by k_uniforms
  uint32_t flags;
  float src_w, src_h;  // texture size when the material textures (camera.go:1560-1566)
  uint32_t light_base, light_count;
  const float* pos;
  const float* nrm;
  const float* col;
  const float* uv;
  const uint16_t* bones;
  const float* weights;
  const float* palette;  // bone_count x 16 floats
  const uint32_t* idx;   // this part's first triangle in the mesh index array (32-bit indices)
  const uint16_t* idx16; // the same for a mesh stored with 16-bit indices (idx is then null)
  const float* ctr;      // Triangle.Center of this part's first triangle, xyz per triangle (mesh.go:1980-1982)
  const float* sk_pos;
  const float* sk_nrm;
  const float* sk_w;     // per-frame shape key weights
  uint32_t shape_keys, mesh_vcount;
  uint32_t vmin, vcount;      // vertex range the part's triangles reference
  uint32_t tcount, vbase;     // triangles; first slot in the vertex scratch
  uint32_t tbase, ntiles;     // first slot in the triangle scratch; tiles of the part
  uint32_t hbase, sblock0;    // histogram offset and first scan block (multi-tile parts)
  uint32_t nsblocks, bone_count;
  uint32_t cam;               // camera (job) of the pair in FrameArgs.fc
  float max_span;             // MaxSpan of the mesh dimensions (mesh.go:41-43), for the point-light skip
  uint32_t light0;            // first FrameArgs.lights entry of the pair's light list
  uint32_t idx_base;          // first index value: 0, or for a later part of a dynamic batch's flush the corners
                              // drawn before it in that flush (k_groups, which also moves out_base past them)
  uint64_t out_base;          // first output vertex
  M4 nrot;                    // rotation of Model.Transform()·ViewMatrix() (model.go:705-708), by k_uniforms
                              // when the pair's camera renders normals
};

// Frame-level light constants (light.go; beginRender results).
struct alignas(16) LightGPU {
  int32_t type;
  float energy;
  float range;
  float range_sq;       // SetRange: range*range (light.go:273-276)
  float color[4];
  float amb[4];         // AmbientLight.result = color·energy (light.go:110-112)
  float fwd[4];         // DirectionalLight.workingForward (light.go:322-324)
  const float* cells;   // LightVolume cell colours, RGBA
  uint32_t ncells;
  int32_t cx, cy, cz;
  float vmin[3], vmax[3], cell[3], offset[3], bias[3];
  float pos[3];         // world position (row 3 of the light's transform), for beginModel
  // LightVolume cell lights (light.go:650-653): [ncells + 1] offsets, then the
  // light indices; index j is entry nest_off + j of the pair's light list
  const uint32_t* cell_lights;
  uint32_t nest_off;
  uint32_t nested;      // 1: one of a volume's cell lights, run only from its cells (never by the list loop)
};

// Per (render pair, light) values from beginModel and the per-model skip,
// written by k_uniforms (frame scratch, never uploaded).
struct alignas(16) PartLight {
  float v[3];           // point: workingPosition; volume per-object mode: cell colour
  uint32_t light;       // index into the frame's LightGPU array
  uint32_t enabled;     // 0: the point-light distance skip removes it (camera.go:1625-1629)
  uint32_t cell_found;  // volume per-object mode: a cell exists at the model position
  uint32_t pad[2];
};

// Camera-level constants (camera.go:1329-1341, 1591-1606).
struct FrameConst {
  M4 vp;
  float campos[4];
  float near_, far_, cam_w, cam_h;
  float half_w, half_h, dmp, spread;
  float snap;
  int32_t perspective, pc_tex, render_depth;
  int32_t fog_active, fog_mode;
  float fog_color[4];
  float fog_range[2];
  int32_t render_normals;   // Camera.RenderNormals: the pairs of this camera fill the normal vertex list
  int32_t pad_fc[3];
  M4 view;                  // ViewMatrix (camera.go:638-648), for the normal list's rotation
};

// One armature whose palette k_palettes computes from bone poses
// (node.go:572-608): palette[b] = inverse_bind·Transform(), Transform() of
// bone b written to FrameArgs.pose_world[world_off + b].
struct PaletteJob {
  const t3d_bone_pose* poses;   // [n], parents first
  const M4* root;               // Transform() of the root bones' parent node, or null
  float* palette;               // [n] x 16 floats
  uint64_t world_off;
  uint32_t n;
  float time;                   // playhead the animation is sampled at
  const t3d_anim_channel* channels;   // animation of the armature (channel b -> bone b), or null
  uint32_t nchannels, pad;
};

// ---- the model-level stage on the device (camera.go:1343-1457, model.go:142-182) ----
//
// With it, the host writes one ModelGPU per drawable model (and the materials
// of its parts) instead of one PartGPU per render pair: k_models tests every
// model's sphere against its camera's frustum and marks each (model, part)
// candidate; k_plist orders the candidates into render pairs (solids in
// model order, then the transparents by render order and distance, stable),
// sizes every pair's slices of the frame arrays and writes the prefix arrays
// the chain's kernels index by; k_uniforms then builds each pair's PartGPU.
// Every later launch is sized for all candidates being drawn; blocks past the
// actual counts (FrameCounts) return at once.

// Resident with each mesh (after its arrays): what the stage reads of it.
struct alignas(16) MeshPartGPU {
  uint32_t triangle_start, triangle_count, vmin, vcount;
};
struct alignas(16) MeshGPU {
  const float* pos;
  const float* nrm;
  const float* col;
  const float* uv;
  const uint16_t* bones;
  const float* weights;
  const uint32_t* idx;
  const uint16_t* idx16;
  const float* ctr;
  const float* sk_pos;
  const float* sk_nrm;
  const MeshPartGPU* parts;
  uint32_t V, K, nparts, pad;
  float dmin[4], dmax[4];
  float max_span;
  uint32_t pad2[3];
};

// Material state of one MeshPart as the stage reads it (t3d_material_desc).
enum : uint32_t {
  MB_PRESENT = 1u << 0, MB_VISIBLE = 1u << 1, MB_BACKFACE = 1u << 2, MB_SHADELESS = 1u << 3,
  MB_BLEND_SOURCE_OVER = 1u << 4, MB_TEXTURE = 1u << 5, MB_BILLBOARD = 1u << 6,
  MB_LOCK_X = 1u << 7, MB_LOCK_Y = 1u << 8, MB_LOCK_Z = 1u << 9, MB_UP_CAMERA_Y = 1u << 10,
  MB_UNBILLBOARDED = 1u << 11,
  MB_FOGLESS = 1u << 12, MB_LINEAR = 1u << 13,   // read by the host only (raster records)
};
struct alignas(16) MaterialGPU {
  float color[4];
  uint32_t bits;            // MB_*
  int32_t transparency_mode;
  int32_t render_order;
  uint32_t modes;           // sort mode | lighting mode << 2 | LightVolume shading mode << 4
  float texture_size[2];
  uint64_t texture;         // Material.Texture handle (with use_texture), for the host's raster records
};

enum : uint32_t {
  MF_FRUSTUM = 1u << 0,     // Model.FrustumCulling
  MF_SKINNED = 1u << 1,
  MF_SHADELESS = 1u << 2,
  MF_DIRECTIONAL = 1u << 3, // a directional light is among the model's lights
  MF_MATERIALS = 1u << 4,   // the model has a material array (MaterialGPU from ModelGPU::mat)
};

struct alignas(16) ModelGPU {
  M4 transform;             // Node.Transform()
  M4 world_rotation;        // Node.WorldRotation()
  float color[4];           // Model.Color
  const MeshGPU* mesh;
  uint32_t mat;             // first MaterialGPU of its parts (MF_MATERIALS)
  uint32_t flags;           // MF_*
  uint32_t light0, light_count;  // the lights its lit pairs see: the job's scene lights or its LightGroup's
  uint32_t cam;             // job of the frame
  uint32_t cand0;           // its first candidate; candidate cand0 + p is its MeshPart p
  uint32_t palette;         // float offset of its bone palette in the table block (skinned)
  uint32_t skw;             // float offset of its shape-key weights in the table block
  uint32_t bone_count;
  uint32_t index;           // its index in its Camera.Render's model list
};

// Per job: the camera's frustum (camera.go:933-983) and what the pairs read of the world.
struct alignas(16) JobGPU {
  float pos[4], fwd[4], up[4], right[4];
  float near_, far_, sf_tang, sf_x;
  float sf_y, aspect, ortho_scale;
  int32_t perspective;
  int32_t lighting_on;      // World != nil && World.LightingOn
  uint32_t cand0, cand1;    // its candidates
  uint32_t pad;
};

// Actual sizes of a frame built on the device (read back with the stats).
enum : uint32_t {
  FC_PAIRS = 0, FC_VCHUNKS, FC_TILES, FC_SBLOCKS, FC_PSCAN, FC_TRANSPARENTS, FC_SUMV, FC_SUMT,
  FC_BYTES_VERTEX_LO, FC_BYTES_VERTEX_HI, FC_BYTES_CULL_LO, FC_BYTES_CULL_HI, FC_BYTES_UV_LO, FC_BYTES_UV_HI,
  FC_WORDS = 16
};

// Per render pair of a device-built frame, read back after the frame: what
// t3d_frame_pairs and the downloads report.
struct alignas(16) PairOut {
  uint32_t model;           // index in its Camera.Render's model list
  uint32_t part;
  uint32_t triangle_count;  // MeshPart.TriangleCount()
  uint32_t tbase;           // first slot in the triangle scratch
  uint64_t out_base;        // first output vertex
  uint32_t triangle_start;  // the part's first mesh triangle
  uint32_t mrec;            // its ModelGPU record
};

struct ModelStage {
  const ModelGPU* models;   // [nmodels], job by job in model order
  const MaterialGPU* mats;
  const JobGPU* jobs;       // [njobs]
  uint32_t nmodels, njobs, ncand, nlights;
  uint32_t* cand;           // [ncand] bit 0: drawn, bit 1: transparent, bits 2-31: its job
  float* cand_dist;         // [ncand] distance² of the candidate's model to its camera
  uint32_t* cand_model;     // [ncand] its ModelGPU record
  int32_t* cand_order;      // [ncand] its Material.RenderOrder
  uint4* cand_cnt;          // [2 * ncand] what its pair would need: vertices, triangles, vertex blocks, tiles;
                            // scan blocks, one-block scan, histogram entries, per-light values
  uint4* cand_bytes;        // [ncand] its algorithmic bytes: vertex pass, cull pass, UVs
  uint4* cand_info;         // [ncand] model index, part, MeshPart triangle count, first mesh triangle
  uint32_t* job_s0;         // [njobs] solid and transparent drawn candidates before each job's first
  uint32_t* job_t0;
  uint32_t* tlist;          // [ncand] drawn transparents, candidate order (jobs contiguous)
  uint32_t* pair_cand;      // [ncand] candidate of each render pair
  uint32_t* job_pairs;      // [njobs + 1] first render pair of each job
  PairOut* pair_out;        // [ncand]
  uint32_t* pc;             // [8][ncand] per pair the exclusive scans of its vertex records (row 0), triangles
                            // (1), histogram entries (6) and per-light values (7); the other rows are unused
                            // (the prefix arrays hold those scans)
  uint32_t* vchunk_prefix;  // [max_pairs + 1] the frame's prefix arrays (FrameArgs), written here
  uint32_t* tile_prefix;
  uint32_t* sblock_prefix;
  uint32_t* pscan_prefix;
  const float* pal_base;    // palettes and shape-key weights of ModelGPU::palette / skw
  const float* skw_base;
  uint32_t vchunk, tile;
  uint32_t max_pairs;       // = ncand
  uint32_t pad;
};

struct FrameArgs {
  const PartGPU* parts;
  uint32_t nparts;
  uint32_t vchunk;              // vertices per vertex-kernel block
  uint32_t tile;                // triangles per tile
  uint32_t xcd_swizzle;         // 1: give each XCD a contiguous range of blocks (see block_index)
  const LightGPU* lights;
  PartLight* plights;
  const uint32_t* vchunk_prefix;  // [P+1]
  const uint32_t* tile_prefix;    // [P+1]
  const uint32_t* sblock_prefix;  // [P+1]
  const uint32_t* pscan_prefix;   // [P+1] one k_scan_part block per multi-tile part with a small matrix
  float4* vclip;                // clip position of vertex record o at vclip[o * vstride]
  float4* vcol;                 // its colour at vcol[o * vstride]
  uint32_t vstride;             // 2: interleaved records; 1: two arrays
  uint32_t pad1;
  float4* vpos;
  float* key;
  uint4* tri_rec;               // [sumT] survivors, triangle order: {vertex-record offsets x3, triangle}
  uint4* tri_vo;                // [sumT] the same records in output-slot order
  uint32_t* stats;              // [P][4]: min bits, max bits, survivors, 0
  uint32_t* tile_count;         // [tiles] survivors of each tile (compacted at the tile's front)
  uint32_t* tile_part;          // [tiles] render pair of each tile (written by k_uniforms)
  uint32_t* vchunk_part;        // [vertex chunks] render pair of each vertex-kernel block (k_uniforms)
  float4* sink;                 // [128] store target of lanes past the end of a chunk (never read)
  uint32_t* hist;
  uint32_t* bsum;
  float4* out_color;            // t3d_vertex as 3 float4
  float4* out_depth;
  uint32_t* out_idx;
  float4* vnrm;                 // [sumV] normal-list colour per vertex record (render_normals cameras only)
  float* vud;                   // [sumV] unbillboarded depth Z per vertex record (PF_UNBB parts only)
  float4* out_normal;           // normal vertices, laid out as out_color (null when no camera renders normals)
  const FrameConst* fc;         // [jobs] device copies, so graph kernel arguments stay fixed across frames
  const PaletteJob* pal_jobs;   // [npal_jobs] armatures posed on the device
  M4* pose_world;               // bone Transform() scratch of the palette jobs
  uint32_t npal_jobs;
  uint32_t ngroups;             // dynamic-batch flushes of more than one render pair
  const uint2* groups;          // [ngroups] {first render pair, render pairs} of each, pairs consecutive
  // a frame whose render pairs the device builds (the model-level stage): its actual sizes (FC_*), and the
  // stage's inputs and scratch; null for a frame built on the host, whose launches are sized exactly
  uint32_t* counts;
  ModelStage ms;
};

// Packing of a range of render pairs' survivors for download: pair j of the
// range copies its first nvert[j] output vertices (and indices) from vertex
// src[j] of the frame outputs to vertex dst[j] of the packed arrays; blocks
// [blk_prefix[j], blk_prefix[j+1]) cover its 3 * nvert float4 words, 256 each.
struct PackArgs {
  const uint32_t* blk_prefix;   // [npairs + 1]
  const uint64_t* src;          // [npairs]
  const uint64_t* dst;          // [npairs]
  const uint32_t* nvert;        // [npairs]
  uint32_t npairs, pad;
  const float4* color;          // frame outputs, t3d_vertex as 3 float4
  const float4* depth;
  const uint32_t* idx;
  float4* pcolor;               // packed (each null when not asked for)
  float4* pdepth;
  uint32_t* pidx;
  const float4* normal;         // frame normal vertices and their packed copy (both null: none packed)
  float4* pnormal;
};

// Raster sink (t3d_raster.hip): the state a flush's draw reads (camera.go:1829-2048).
enum : uint32_t {
  RF_LINEAR = 1u << 0,      // TextureFilterMode linear (base3d.kage bilinearFilter)
  RF_CLIP = 1u << 1,        // TransparencyModeAlphaClip: the clip-alpha depth shader (camera.go:394-478)
  RF_FOGLESS = 1u << 2,     // Material.Fogless
  RF_COMPOSITE = 1u << 3,   // not transparent: its depth intermediate is drawn over the result depth
  RF_BAD_BLEND = 1u << 4,   // a blend other than SourceOver: the raster sink refuses the frame
  RF_NO_TEXTURE = 1u << 5,  // the material names a texture handle the context does not hold
};
struct RasterFlush {
  const uint8_t* tex;       // RGBA8 premultiplied, or the default 1x1 white image (tetra3d.go:12, 38)
  uint32_t tw, th;
  uint32_t flags;
  uint32_t pad;
};
// Camera and world uniforms of the draws (camera.go:1002-1013, 1956-1977; world.go:96-110).
struct RasterFrame {
  float fog[4];             // FogColor RGB, FogMode (-1 when the world's fog is off)
  float fog_range[2];
  float dither;             // DitheredFogSize
  int32_t curve;            // FogCurve
  float clear[4];           // ClearColor, straight alpha (transparent black without a world)
  int32_t width, height;
  int32_t render_depth, pc_tex;
};
// Per render pair, for the raster kernels: its first drawn triangle among all
// the frame's, where its corners are, and its flush.
struct RasterPair {
  uint32_t tri0;            // prefix of survivors over the render pairs before it
  uint32_t flush;
  uint64_t vertex;          // first output vertex of its survivors
};

// Kernel launchers (t3d_kernels.hip).  blk0 offsets blockIdx so the plain
// form can launch one part's blocks at a time.
hipError_t launch_palettes(const FrameArgs& a, hipStream_t s);
hipError_t launch_uniforms(const FrameArgs& a, uint32_t part0, uint32_t nparts, hipStream_t s);
hipError_t launch_vertex(const FrameArgs& a, uint32_t blk0, uint32_t nblk, uint32_t threads, hipStream_t s);
hipError_t launch_cull(const FrameArgs& a, uint32_t blk0, uint32_t nblk, uint32_t threads, hipStream_t s);
hipError_t launch_hist(const FrameArgs& a, uint32_t blk0, uint32_t nblk, hipStream_t s);
hipError_t launch_scan(const FrameArgs& a, uint32_t sblk0, uint32_t nsblk, uint32_t part0, uint32_t nparts,
                       hipStream_t s);
hipError_t launch_scan_part(const FrameArgs& a, uint32_t blk0, uint32_t nblk, hipStream_t s);
hipError_t launch_scatter(const FrameArgs& a, uint32_t blk0, uint32_t nblk, uint32_t waves, hipStream_t s);
hipError_t launch_assemble(const FrameArgs& a, uint32_t blk0, uint32_t nblk, uint32_t mode, uint32_t split,
                           hipStream_t s);
hipError_t launch_pack(const PackArgs& a, uint32_t nblocks, hipStream_t s);
hipError_t launch_normals(const FrameArgs& a, uint32_t blk0, uint32_t nblk, hipStream_t s);
hipError_t launch_groups(const FrameArgs& a, uint32_t part0, uint32_t part1, hipStream_t s);
// The model-level stage of a device-built frame (t3d_models.hip): k_models, then k_plist.
hipError_t launch_model_stage(const FrameArgs& a, hipStream_t s);
hipError_t launch_copy(const void* src, void* dst, uint64_t bytes, hipStream_t s);
hipError_t launch_fill(void* dst, uint64_t bytes, uint32_t word, hipStream_t s);
hipError_t launch_fp_selftest(const float* a, const float* b, float* out, uint32_t n, hipStream_t s);

}  // namespace t3d
