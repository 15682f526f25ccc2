// t3d_runtime.h — host-side runtime state shared by the render path
// (t3d_host.cpp) and the serving loop (t3d_serve.cpp).
#pragma once

#include <hip/hip_runtime.h>

#include <chrono>
#include <map>
#include <memory>
#include <mutex>
#include <string>
#include <vector>

#include "t3d_internal.h"

namespace t3d {

// A device allocation accounted against the context budget.
struct DevBuf {
  void* ptr = nullptr;
  uint64_t bytes = 0;
};

// A resident mesh (mesh.go:260-302) and its MeshPart ranges.
struct Mesh {
  uint64_t handle = 0;
  DevBuf buf;
  const float* pos = nullptr;
  const float* nrm = nullptr;
  const float* col = nullptr;
  const float* uv = nullptr;
  const uint16_t* bones = nullptr;
  const float* weights = nullptr;
  const float* sk_pos = nullptr;
  const float* sk_nrm = nullptr;
  const uint32_t* idx = nullptr;   // triangle indices, uint32 ...
  const uint16_t* idx16 = nullptr; // ... or uint16 when the mesh has at most 65,536 vertices
  const float* ctr = nullptr;      // Triangle.Center per triangle, xyz (mesh.go:1980-1982, kept as AddTriangles does)
  uint32_t V = 0, T = 0, K = 0;
  std::vector<t3d_part_range> parts;
  std::vector<uint32_t> part_vmin, part_vcount;
  float dmin[3], dmax[3];
  float max_span = 0;              // MaxSpan(Dimensions) (mesh.go:41-43)
  uint32_t max_bone_index = 0;
  bool has_skin = false;
  int refs = 0;                    // submitted jobs referencing the mesh
  const MeshGPU* gpu = nullptr;    // its record for the device model stage (in buf, after the arrays)
  // what all its parts would need if every one were drawn, for launch shape (vchunk, tile) (device model stage)
  uint32_t wc_vchunk = 0, wc_tile = 0;
  uint32_t wc_vchunks = 0, wc_tiles = 0, wc_sblocks = 0, wc_pscan = 0;
  uint64_t wc_hist = 0, wc_V = 0, wc_T = 0;
};

struct Volume {
  uint64_t handle = 0;
  DevBuf buf;
  uint32_t n = 0;
  std::vector<float> host;         // for per-object lookups on the host
  int refs = 0;                    // submitted jobs referencing the volume
  DevBuf lists;                    // cell light lists: [n + 1] offsets then the light indices (empty: none)
  uint32_t lists_need = 0;         // 1 + the largest light index the lists name (0: no lists)
};

struct Animation {
  uint64_t handle = 0;
  DevBuf buf;                       // channels (device pointers) followed by their times and values
  uint32_t nchannels = 0;
  const t3d_anim_channel* channels = nullptr;   // device address
  int refs = 0;                     // submitted jobs referencing the animation
};

struct Texture {
  uint64_t handle = 0;
  DevBuf buf;                       // RGBA8, premultiplied, row-major
  uint32_t w = 0, h = 0;
};

}  // namespace t3d

struct t3d_frame {
  // device tables and scratch
  t3d::DevBuf tables, scratch, outputs;
  uint64_t tables_cap = 0, scratch_cap = 0, outputs_cap = 0;
  // download staging of t3d_frame_download: the survivors packed back to back, so the copy to the host is one
  // transfer per stream (grown on demand, kept with the frame)
  t3d::DevBuf pack;
  uint64_t pack_cap = 0;
  void* pinned = nullptr;          // host staging of the tables
  uint64_t pinned_cap = 0;
  uint32_t* pinned_stats = nullptr;
  uint64_t pinned_stats_cap = 0;
  // a render placed in external memory (t3d_render_noalloc): pageable host staging of the tables and of the
  // stats readback, and where the last render's stats were read back to
  std::vector<uint8_t> host_tables;
  std::vector<uint32_t> host_stats;
  const uint32_t* stats_host = nullptr;
  uint8_t* tables_dev = nullptr;   // device table block of the last build (own or placed)
  bool any_normals = false;        // the last build fills the normal vertex list (Camera.RenderNormals)
  t3d::FrameArgs args{};
  // per-render bookkeeping
  std::vector<t3d_pair_result> pairs;
  std::vector<t3d::PartGPU> parts;
  std::vector<uint32_t> part_first_tri;  // mesh triangle_start per pair
  std::vector<uint64_t> bytes_vertex, bytes_cull, bytes_uv;  // algorithmic bytes per part
  uint64_t bytes_palettes = 0;                               // of k_palettes, all armatures
  std::vector<t3d::FrameConst> fcs;      // one camera per job of the frame
  std::vector<uint32_t> job_pair_begin;  // [jobs + 1] first render pair of each job
  // Flushes (the render pairs the caller sees): with no dynamic batch in the frame, one per render pair and
  // `pairs` itself; with one, `flush` holds a record per flush, flush_begin [flushes + 1] its first render pair,
  // job_flush_begin [jobs + 1] each job's first flush, and groups the batches of more than one render pair.
  bool grouped = false;
  std::vector<t3d_pair_result> flush;
  std::vector<uint32_t> flush_begin;
  std::vector<uint32_t> job_flush_begin;
  std::vector<uint2> groups;
  const std::vector<t3d_pair_result>& out_pairs() const { return grouped ? flush : pairs; }
  // raster sink: per flush the state its draw reads (material texture, filter, transparency), per frame the
  // camera and world uniforms, and the images and working buffers of t3d_frame_rasterize
  std::vector<t3d::RasterFlush> rflush;
  t3d::RasterFrame rframe{};
  std::vector<uint64_t> rflush_tex;   // texture handle of each flush (0: the default image), resolved again at draw time
  t3d::DevBuf raster_buf, raster_buf2, raster_buf3, images;
  uint64_t raster_cap = 0, raster_cap2 = 0, raster_cap3 = 0, images_cap = 0;
  uint8_t* image_color = nullptr;   // device images of the last rasterize (null before)
  uint8_t* image_depth = nullptr;
  int32_t image_w = 0, image_h = 0;
  uint32_t job_out_begin(uint32_t k) const { return grouped ? job_flush_begin[k] : job_pair_begin[k]; }
  std::vector<uint32_t> pair_tbase;      // first triangle-scratch slot of each render pair
  // the model-level stage ran on the device: the chain is launched for max_pairs candidate pairs, and the
  // render pairs, their offsets and the jobs' first pairs are read back with the stats (rb: counts, job
  // pairs, PairOut per pair, stats per pair, at rb_* byte offsets)
  bool device_stage = false;
  uint32_t max_pairs = 0, stage_jobs = 0;
  uint64_t rb_bytes = 0, rb_jobs = 0, rb_pairs = 0, rb_stats = 0;
  uint8_t* rb_dev = nullptr;
  uint8_t* pinned_rb = nullptr;
  uint64_t pinned_rb_cap = 0;
  const t3d::ModelGPU* stage_models = nullptr;     // host copies in the table staging, for the raster records
  const t3d::MaterialGPU* stage_mats = nullptr;
  uint64_t table_bytes = 0;
  uint64_t total_triangles = 0;
  uint32_t nparts = 0;
  uint32_t total_vchunks = 0, total_tiles = 0, total_sblocks = 0;
  std::vector<uint32_t> vchunk_prefix, tile_prefix, sblock_prefix, pscan_prefix;
  uint32_t total_pscan = 0;
  uint32_t vertex_threads = 256;
  uint64_t total_survivors = 0;
  double last_ms = 0;
  // graph cache of the batched chain
  hipGraphExec_t graph_exec = nullptr;
  std::vector<uint8_t> graph_key;
  // asynchronous renders: tables uploaded / whole frame done, and whether a
  // launch is outstanding
  hipEvent_t tables_ready = nullptr, done = nullptr;
  hipStream_t stream = nullptr;    // the frame's own chain stream (null: the caller's)
  bool inflight = false;
  std::chrono::steady_clock::time_point submitted;
};

struct t3d_ctx {
  int device = 0;
  hipStream_t stream = nullptr;        // kernel chains of t3d_render / t3d_render_async
  hipStream_t copy_stream = nullptr;   // table uploads, so they overlap the previous frame's kernels
  uint64_t budget = 0, used = 0;
  uint64_t next_handle = 1;
  std::map<uint64_t, std::unique_ptr<t3d::Mesh>> meshes;
  std::map<uint64_t, std::unique_ptr<t3d::Volume>> volumes;
  std::map<uint64_t, std::unique_ptr<t3d::Animation>> animations;
  std::map<uint64_t, std::unique_ptr<t3d::Texture>> textures;
  std::string err;
  t3d_tuning tuning{256, 1, 2048, 4, 1, 2, 8, 1, 0, 256, 1};
  // Guards resources, the budget and err between callers and the serving
  // thread; frame builds hold it, device work runs outside it.
  std::mutex mu;
  // nonzero: fill every frame's outputs with this word before its render
  uint32_t poison = 0;
  // this device's copy of the static scratch array of t3d_render_noalloc (looked up at context creation)
  void* noalloc_scratch = nullptr;
  // profiling (synchronous render path)
  bool profiling = false;
  struct Prof {
    double ms = 0;
    uint32_t launches = 0;
    uint64_t bytes = 0;
  };
  std::map<std::string, Prof> prof;
};

namespace t3d {

t3d_status fail(t3d_ctx* c, t3d_status st, const char* fmt, ...);
// Device address of `device`'s copy of the no-allocation scratch array (t3d_noalloc.hip); the device must be
// current.  Looked up once per device.
hipError_t noalloc_scratch_address(int device, void** out);
t3d_status dev_alloc(t3d_ctx* c, DevBuf& b, uint64_t bytes);
// Grow `b` to at least `bytes` (contents discarded); the caller holds ctx->mu.
t3d_status ensure(t3d_ctx* c, DevBuf& b, uint64_t& cap, uint64_t bytes);
void dev_free(t3d_ctx* c, DevBuf& b);
void frame_release(t3d_ctx* c, t3d_frame* f);
// Device bytes a frame holds between renders (tables, scratch, outputs, download packing).
inline uint64_t frame_capacity(const t3d_frame* f) {
  return f->tables.bytes + f->scratch.bytes + f->outputs.bytes + f->pack.bytes + f->raster_buf.bytes +
         f->raster_buf2.bytes + f->raster_buf3.bytes + f->images.bytes;
}

// Every refusal of a Camera.Render descriptor that does not depend on the frustum test (caller holds ctx->mu);
// frame_build runs it for each descriptor, the serving form at submit.
t3d_status validate_desc(t3d_ctx* c, const t3d_render_desc* desc);

// Device memory a frame uses without allocating (t3d_render_noalloc): its
// tables and scratch are carved from `scratch`, its outputs are the caller's.
struct Placement {
  uint8_t* scratch = nullptr;
  uint64_t scratch_bytes = 0;
  t3d_vertex* color = nullptr;
  t3d_vertex* depth = nullptr;
  uint32_t* idx = nullptr;
  uint64_t capacity = 0;         // vertices each output array holds
};

// One frame of one or more Camera.Render calls: build (host work, caller
// holds ctx->mu), launch (enqueue on `stream`), collect (after completion).
// With `place`, nothing is allocated: the frame's device memory is the
// placement's, its table staging pageable host memory.
// With device_stage, a frame that allows it (and whose tuning asks for it) has its model-level stage run on the
// device: frustum tests, render pairs and their slices are then known only after the frame (frame_collect).
t3d_status frame_build(t3d_ctx* c, t3d_frame* f, const t3d_render_desc* const* descs, uint32_t ndesc,
                       const Placement* place = nullptr, bool device_stage = false);
// Runs a frame built with a placement on `stream` and waits for it: table
// upload from pageable memory, the chain (per pair for the plain form, one
// chain for the batched form; never a graph), stats readback.  Calls no
// allocating HIP routine.
t3d_status frame_run_placed(t3d_ctx* c, t3d_frame* f, hipStream_t stream, int32_t exec_mode);
// The table upload goes on `copy_stream` (may equal `stream`); `stream` waits
// for it, runs the chain and the survivor readback, then records f->done.
t3d_status frame_launch(t3d_ctx* c, t3d_frame* f, hipStream_t stream, hipStream_t copy_stream, int32_t exec_mode,
                        bool allow_graph, bool profile);
void frame_collect(t3d_ctx* c, t3d_frame* f, bool profile);
// Device bytes frame_download_range needs in `pack` for pairs [p0, p1) (0 when none survive).
uint64_t download_pack_bytes(const t3d_frame* f, uint32_t p0, uint32_t p1, bool normals = false);
// Copy pairs [p0, p1) of a frame, concatenated, to host memory (any output may be null).  Vertices and
// indices are first packed on the device into `pack`, so they cross to the host in one copy per stream.
// Without err_out, `pack` is grown here when too small (needs ctx->mu) and errors go to ctx->err; with it,
// the caller has grown `pack` to download_pack_bytes already and the message goes to *err_out (no ctx state
// is touched).  Returns only after every copy it enqueued has finished.
t3d_status frame_download_range(t3d_ctx* c, const t3d_frame* f, uint32_t p0, uint32_t p1, hipStream_t stream,
                                t3d_vertex* cv, t3d_vertex* dv, uint32_t* idx, uint32_t* tri_ids, t3d_vertex* nv,
                                t3d::DevBuf* pack = nullptr, uint64_t* pack_cap = nullptr,
                                std::string* err_out = nullptr);
// Frame memory a job may need (scratch + outputs + tables), before frustum culling.
uint64_t estimate_job_bytes(t3d_ctx* c, const t3d_render_desc* desc);

}  // namespace t3d
