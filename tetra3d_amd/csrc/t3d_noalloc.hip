// t3d_noalloc.hip — Camera.Render into caller-owned outputs with no device
// allocation (t3d_render_noalloc).
//
// The frame's tables (per-pair uniforms, lights, block prefixes, palettes)
// and its scratch (vertex records, depth keys, survivor records, histograms,
// per-pair stats) come from one file-scope device array of
// T3D_NOALLOC_SCRATCH_BYTES per GPU, which each GPU holds until the process
// ends.  The host side reaches it only through the address
// hipGetSymbolAddress returns for the current device: the array's name in
// host code is a host-side stand-in, not a device address.
//
// Every call on one GPU shares that GPU's array and leaves its contents
// behind; nothing in it needs to start at zero (k_uniforms resets each pair's
// depth range and survivor count, every other slice is written before it is
// read).  Calls that overlap on one GPU, from any context, race on it.
#include <hip/hip_runtime.h>

#include <mutex>

#include "t3d_runtime.h"

namespace t3d {

static __device__ __align__(256) uint8_t scratch[T3D_NOALLOC_SCRATCH_BYTES];

namespace {

constexpr int kMaxDevices = 64;
std::mutex g_mu;
void* g_ptr[kMaxDevices] = {};

}  // namespace

hipError_t noalloc_scratch_address(int device, void** out) {
  *out = nullptr;
  if (device < 0 || device >= kMaxDevices) return hipErrorInvalidDevice;
  std::lock_guard<std::mutex> lk(g_mu);
  if (!g_ptr[device]) {
    int cur = -1;
    hipError_t e = hipGetDevice(&cur);
    if (e != hipSuccess) return e;
    if (cur != device) return hipErrorInvalidDevice;   // the caller makes `device` current
    void* p = nullptr;
    e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess) return e;
    g_ptr[device] = p;
  }
  *out = g_ptr[device];
  return hipSuccess;
}

}  // namespace t3d

using namespace t3d;

extern "C" t3d_status t3d_render_noalloc(t3d_ctx* c, t3d_frame* f, const t3d_render_desc* desc,
                                         t3d_vertex* color_vertices, t3d_vertex* depth_vertices, uint32_t* indices,
                                         uint64_t vertex_capacity) {
  if (!c || !f || !desc) return fail(c, T3D_ERR_INVALID_ARGUMENT, "render_noalloc: null argument");
  if (!color_vertices || !depth_vertices || !indices)
    return fail(c, T3D_ERR_INVALID_ARGUMENT, "render_noalloc: null output array");
  std::lock_guard<std::mutex> lk(c->mu);
  if (!c->noalloc_scratch) return fail(c, T3D_ERR_DEVICE, "render_noalloc: no scratch address for device %d", c->device);
  if (f->inflight) {   // an asynchronous render of this frame: finish it, its results are replaced
    f->inflight = false;
    hipError_t e = hipEventSynchronize(f->done);
    if (e != hipSuccess) return fail(c, T3D_ERR_DEVICE, "render_noalloc: %s", hipGetErrorString(e));
  }
  const auto wall0 = std::chrono::steady_clock::now();
  Placement pl;
  pl.scratch = reinterpret_cast<uint8_t*>(c->noalloc_scratch);
  pl.scratch_bytes = T3D_NOALLOC_SCRATCH_BYTES;
  pl.color = color_vertices;
  pl.depth = depth_vertices;
  pl.idx = indices;
  pl.capacity = vertex_capacity;
  t3d_status st = frame_build(c, f, &desc, 1, &pl);
  if (st) {
    f->nparts = 0;   // nothing of a refused frame is published
    f->pairs.clear();
    f->flush.clear();
    f->grouped = false;
    return st;
  }
  st = frame_run_placed(c, f, c->stream, desc->exec_mode);
  if (st) return st;
  f->last_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - wall0).count();
  return T3D_OK;
}
