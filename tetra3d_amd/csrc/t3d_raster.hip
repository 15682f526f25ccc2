// t3d_raster.hip — the raster sink: each flush of a rendered frame drawn into
// the camera's result textures on the GPU, as Camera.Render's flush does
// through ebiten (camera.go:1829-2048) with the Kage shaders (camera.go:350-
// 531, shaders/base3d.kage:61-153).  The colour texture (and, with
// RenderDepth, the depth texture) stay in HBM; a host that needs the picture
// copies 4 bytes per pixel instead of the frame's vertex lists.
//
// ebiten hands the vertex lists to the GPU's fixed-function rasteriser,
// which is not part of the reference; its rules are fixed here:
//  - images are RGBA8 with premultiplied alpha; a float x is stored as
//    floor(clamp(x, 0, 1) * 255 + 0.5);
//  - pixel (x, y) is sampled at (x + 0.5, y + 0.5); a triangle covers it when
//    its three edge functions, in double precision from each edge's endpoints
//    taken smaller (y, x) first (so a shared edge gives exactly opposite
//    values), have the sign of the area, a zero counting on a top or left edge;
//  - attributes are interpolated linearly in screen space with barycentric
//    weights edge / area (double precision), then rounded to float32;
//  - triangles of one draw go in index order; blending is SourceOver,
//    src + dst * (1 - src.a), float32 per channel.
//
// Work split (every step a parallel kernel): a setup pass writes one
// 160-byte record per drawn triangle (corners, the attributes the draws read,
// area, pixel box, flush) and the area of its pixel box; a scan of the areas
// numbers every box pixel ("slot"), and a table gives each 256 slots the
// triangles of their first and last; one thread per slot tests coverage, a
// wavefront's answers one 64-bit mask; a scan of the mask counts gives each
// covered slot its fragment, written as (pixel, triangle and place in its
// box); a stable radix sort by pixel leaves each pixel's fragments contiguous
// and in draw order, where they are shaded (interpolation, texture, clip
// alpha, base3d's colour before fog) into 32-byte fragments; one thread per
// pixel replays them flush by flush: the depth pass into the intermediate,
// the composite for opaque pairs, the colour pass with fog.
// Every fragment of a frame (about 21 million on mixed-512 at 1920 x 1080,
// ten per pixel) is shaded with full SIMD lanes, and no pixel waits on
// another's triangles.
#include <hip/hip_runtime.h>
#include <hipcub/hipcub.hpp>

#include <cmath>
#include <cstring>

#include "t3d_internal.h"
#include "t3d_runtime.h"

namespace t3d {
namespace {

struct STri;

struct Frag;

struct RArgs {
  const RasterPair* pairs;        // render pairs with survivors, in draw order
  uint32_t npairs;
  uint32_t ntri;                  // drawn triangles of the frame
  const RasterFlush* flushes;
  const float4* color;            // frame outputs, t3d_vertex as 3 float4
  const float4* depth;
  RasterFrame rf;
  STri* recs;                     // [ntri] the drawn triangles' records
  uint64_t* counts;               // [ntri + 1] pixels of each triangle's box (its fragment slots)
  const uint64_t* offsets;        // [ntri + 1] exclusive scan of counts
  uint64_t nslots;
  uint2* btri;                    // [nslot blocks] triangles of each 256 slots' first and last slot
  uint64_t* masks;                // [nwaves] per 64 slots: which cover their pixel
  uint32_t* wcount;               // [nwaves + 1] covered slots of each 64 (and a closing 0)
  const uint32_t* wofs;           // [nwaves + 1] exclusive scan of wcount: each 64's first fragment
  uint32_t nwaves;
  uint32_t nfrags;
  uint32_t* keys;                 // [nfrags] pixel of each fragment, draw order
  uint64_t* vals;                 // [nfrags] its triangle (high word) and place in the triangle's box
  const uint32_t* skeys;          // both sorted by pixel (stable: draw order within a pixel)
  const uint64_t* svals;
  Frag* sfrags;                   // [nfrags] the fragments shaded in that order
  uint32_t* pix_begin;            // [W * H] range of each pixel in the sorted fragments
  uint32_t* pix_end;
  uint32_t* seg_start;            // [nfrags + 1] 1 where a (pixel, flush) run of the sorted fragments starts
  const uint32_t* seg_id;         // [nfrags + 1] exclusive scan of seg_start: the run of each fragment (+1 at a start)
  uint32_t* seg_end;              // [runs] one past the last fragment of each run
  uint32_t* img_color;            // RGBA8 packed, r in the low byte
  uint32_t* img_depth;
};

// One covered fragment, shaded: what the per-pixel resolve needs.
struct alignas(16) Frag {
  uint32_t flush;
  float d;          // the depth list's interpolated ColorR
  float clip_a;     // AlphaClip: the texel alpha (< 0: discarded by the clip shader); otherwise 1
  uint32_t pad;
  float col[4];     // base3d's colorTex before fog (RenderDepth), or the plain draw's source colour
};


// ---- the arithmetic of the draw (float32 shaders, double coverage) ----------

__device__ __forceinline__ uint32_t q8(float x) {
  if (!(x > 0.0f)) x = 0.0f;
  if (x > 1.0f) x = 1.0f;
  return (uint32_t)floorf(x * 255.0f + 0.5f);
}
__device__ __forceinline__ float fractf_(float x) { return x - floorf(x); }
__device__ __forceinline__ float modf_(float x, float y) { return x - y * floorf(x / y); }
__device__ __forceinline__ float mixf_(float a, float b, float t) { return a * (1.0f - t) + b * t; }
__device__ __forceinline__ float chan(uint32_t px, int k) { return (float)((px >> (8 * k)) & 255u) / 255.0f; }
__device__ __forceinline__ uint32_t pack(const float o[4]) {
  return q8(o[0]) | (q8(o[1]) << 8) | (q8(o[2]) << 16) | (q8(o[3]) << 24);
}
__device__ __forceinline__ uint32_t source_over(uint32_t dst, const float src[4]) {
  float o[4];
  for (int k = 0; k < 4; k++) o[k] = src[k] + chan(dst, k) * (1.0f - src[3]);
  return pack(o);
}
// encodeDepth / decodeDepth (camera.go:354-364)
__device__ __forceinline__ void encode_depth(float d, float out[4]) {
  out[0] = floorf(d * 255.0f) / 255.0f;
  out[1] = floorf(fractf_(d * 255.0f) * 255.0f) / 255.0f;
  out[2] = fractf_(d * 255.0f * 255.0f);
  out[3] = 1.0f;
}
__device__ __forceinline__ float decode_depth(uint32_t px) {
  return chan(px, 0) + (chan(px, 1) / 255.0f) + (chan(px, 2) / 65025.0f);
}

// imageSrcNUnsafeAt for an unatlased image: the texel containing (x, y); a
// null image is the default 1x1 white one (tetra3d.go:12, 38)
__device__ __forceinline__ void texel_at(const uint8_t* img, uint32_t w, uint32_t h, float x, float y, float out[4]) {
  if (!img) {
    for (int k = 0; k < 4; k++) out[k] = 1.0f;
    return;
  }
  const float fx = floorf(x), fy = floorf(y);
  const int ix = fx != fx ? 0 : (fx < 0.0f ? 0 : (fx >= (float)w ? (int)w - 1 : (int)fx));
  const int iy = fy != fy ? 0 : (fy < 0.0f ? 0 : (fy >= (float)h ? (int)h - 1 : (int)fy));
  const uchar4 p = reinterpret_cast<const uchar4*>(img)[(size_t)iy * w + ix];
  out[0] = (float)p.x / 255.0f;
  out[1] = (float)p.y / 255.0f;
  out[2] = (float)p.z / 255.0f;
  out[3] = (float)p.w / 255.0f;
}
// nearest, or base3d's bilinearFilter (base3d.kage:33-50)
__device__ void sample(const RasterFlush& fl, float x, float y, float out[4]) {
  if (!(fl.flags & RF_LINEAR)) {
    texel_at(fl.tex, fl.tw, fl.th, x, y, out);
    return;
  }
  const float w = (float)fl.tw, h = (float)fl.th;
  const float p0x = modf_(x - 1.0f / 2.0f, w), p0y = modf_(y - 1.0f / 2.0f, h);
  const float p1x = modf_(x + 1.0f / 2.0f, w), p1y = modf_(y + 1.0f / 2.0f, h);
  float c0[4], c1[4], c2[4], c3[4];
  texel_at(fl.tex, fl.tw, fl.th, p0x, p0y, c0);
  texel_at(fl.tex, fl.tw, fl.th, p1x, p0y, c1);
  texel_at(fl.tex, fl.tw, fl.th, p0x, p1y, c2);
  texel_at(fl.tex, fl.tw, fl.th, p1x, p1y, c3);
  const float rx = fractf_(p1x), ry = fractf_(p1y);
  for (int k = 0; k < 4; k++) out[k] = mixf_(mixf_(c0[k], c1[k], rx), mixf_(c2[k], c3[k], rx), ry);
}

__device__ __forceinline__ double edge_fn(float ax, float ay, float bx, float by, double cx, double cy) {
  const bool swap = (ay > by) || (ay == by && ax > bx);
  const double px = swap ? bx : ax, py = swap ? by : ay, qx = swap ? ax : bx, qy = swap ? ay : by;
  const double e = (qx - px) * (cy - py) - (qy - py) * (cx - px);
  return swap ? -e : e;
}
__device__ __forceinline__ bool top_left(float ax, float ay, float bx, float by, double s) {
  const double dx = s * ((double)bx - (double)ax), dy = s * ((double)by - (double)ay);
  return (dy == 0.0 && dx > 0.0) || dy < 0.0;
}

// A drawn triangle's record (written once by k_rsetup, 160 bytes): corner
// positions and the attributes the draws read.
struct alignas(16) STri {
  float x[3], y[3];          // DstX, DstY
  float sx[3], sy[3], c0[3]; // SrcX, SrcY, Custom0 (the same on both lists)
  float r[3], g[3], b[3], a[3];
  float dz[3];               // the depth list's ColorR
  double area;               // orientation: its sign
  int bx0, bx1, by0, by1;    // pixel box on the image; bx0 > bx1: nothing to draw
  uint32_t flush;
  uint32_t pad[3];
};
static_assert(sizeof(STri) == 160, "raster record layout");

// Area, orientation and the box of pixels whose centres the corners span.
__device__ __forceinline__ bool setup(STri& t, int W, int H) {
  t.area = edge_fn(t.x[0], t.y[0], t.x[1], t.y[1], (double)t.x[2], (double)t.y[2]);
  t.bx0 = 1;
  t.bx1 = 0;
  t.by0 = 1;
  t.by1 = 0;
  if (!(t.area != 0.0) || t.area != t.area) return false;
  double mnx = t.x[0], mxx = t.x[0], mny = t.y[0], mxy = t.y[0];
  for (int k = 1; k < 3; k++) {
    mnx = t.x[k] < mnx ? t.x[k] : mnx;
    mxx = t.x[k] > mxx ? t.x[k] : mxx;
    mny = t.y[k] < mny ? t.y[k] : mny;
    mxy = t.y[k] > mxy ? t.y[k] : mxy;
  }
  const double fx0 = ceil(mnx - 0.5), fx1 = floor(mxx - 0.5), fy0 = ceil(mny - 0.5), fy1 = floor(mxy - 0.5);
  if (!(fx0 <= fx1) || !(fy0 <= fy1) || fx1 < 0 || fy1 < 0 || fx0 > W - 1 || fy0 > H - 1) return false;
  t.bx0 = fx0 < 0 ? 0 : (int)fx0;
  t.by0 = fy0 < 0 ? 0 : (int)fy0;
  t.bx1 = fx1 > W - 1 ? W - 1 : (int)fx1;
  t.by1 = fy1 > H - 1 ? H - 1 : (int)fy1;
  return true;
}

// coverage alone (no barycentric weights): the same tests as cover()
__device__ __forceinline__ bool cover_test(const STri& t, int x, int y) {
  if (x < t.bx0 || x > t.bx1 || y < t.by0 || y > t.by1) return false;
  const double cx = x + 0.5, cy = y + 0.5, s = t.area > 0 ? 1.0 : -1.0;
  const double e12 = edge_fn(t.x[1], t.y[1], t.x[2], t.y[2], cx, cy);
  if (!(s * e12 > 0 || (e12 == 0 && top_left(t.x[1], t.y[1], t.x[2], t.y[2], s)))) return false;
  const double e20 = edge_fn(t.x[2], t.y[2], t.x[0], t.y[0], cx, cy);
  if (!(s * e20 > 0 || (e20 == 0 && top_left(t.x[2], t.y[2], t.x[0], t.y[0], s)))) return false;
  const double e01 = edge_fn(t.x[0], t.y[0], t.x[1], t.y[1], cx, cy);
  return s * e01 > 0 || (e01 == 0 && top_left(t.x[0], t.y[0], t.x[1], t.y[1], s));
}

__device__ __forceinline__ bool cover(const STri& t, int x, int y, double l[3]) {
  if (x < t.bx0 || x > t.bx1 || y < t.by0 || y > t.by1) return false;
  const double cx = x + 0.5, cy = y + 0.5, s = t.area > 0 ? 1.0 : -1.0;
  const double e12 = edge_fn(t.x[1], t.y[1], t.x[2], t.y[2], cx, cy);
  const double e20 = edge_fn(t.x[2], t.y[2], t.x[0], t.y[0], cx, cy);
  const double e01 = edge_fn(t.x[0], t.y[0], t.x[1], t.y[1], cx, cy);
  if (!(s * e12 > 0 || (e12 == 0 && top_left(t.x[1], t.y[1], t.x[2], t.y[2], s)))) return false;
  if (!(s * e20 > 0 || (e20 == 0 && top_left(t.x[2], t.y[2], t.x[0], t.y[0], s)))) return false;
  if (!(s * e01 > 0 || (e01 == 0 && top_left(t.x[0], t.y[0], t.x[1], t.y[1], s)))) return false;
  l[0] = e12 / t.area;
  l[1] = e20 / t.area;
  l[2] = e01 / t.area;
  return true;
}
__device__ __forceinline__ float interp(const double l[3], const float a[3]) {
  return (float)(l[0] * (double)a[0] + l[1] * (double)a[1] + l[2] * (double)a[2]);
}

// The render pair holding drawn triangle t (pairs with survivors only).
__device__ __forceinline__ uint32_t pair_of(const RArgs& a, uint32_t t) {
  uint32_t lo = 0, hi = a.npairs;   // largest p with tri0 <= t
  while (hi - lo > 1) {
    const uint32_t mid = (lo + hi) / 2;
    if (a.pairs[mid].tri0 <= t) lo = mid;
    else hi = mid;
  }
  return lo;
}

__device__ void load_tri(const RArgs& a, uint32_t t, STri& s) {
  const RasterPair rp = a.pairs[pair_of(a, t)];
  const uint64_t v0 = rp.vertex + 3ull * (t - rp.tri0);
  for (int k = 0; k < 3; k++) {
    const float4 p = a.color[3 * (v0 + k) + 0];
    const float4 c = a.color[3 * (v0 + k) + 1];
    const float4 q = a.color[3 * (v0 + k) + 2];
    s.x[k] = p.x;
    s.y[k] = p.y;
    s.sx[k] = p.z;
    s.sy[k] = p.w;
    s.r[k] = c.x;
    s.g[k] = c.y;
    s.b[k] = c.z;
    s.a[k] = c.w;
    s.c0[k] = q.x;
    s.dz[k] = a.rf.render_depth ? a.depth[3 * (v0 + k) + 1].x : 0.0f;
  }
  s.flush = rp.flush;
}

// ---- shading ---------------------------------------------------------------------

// base3d.kage Fragment (61-153), split in two: the texture and vertex colour
// (colorTex after colorTex.rgb *= color.a), computed with the fragment, and the
// fog, which reads the flush's final depth intermediate at the pixel
__device__ void base3d_color(const RasterFlush& fl, const RasterFrame& rf, const float vc[4], float sx, float sy,
                             float c0, float out[4]) {
  float tx = sx, ty = sy;
  if (rf.pc_tex > 0) {                                              // tx *= 1.0 / custom.x
    const float r = 1.0f / c0;
    tx *= r;
    ty *= r;
  }
  tx = modf_(tx, (float)fl.tw);                                     // mod(tx, srcSize)
  ty = modf_(ty, (float)fl.th);
  float tex[4];
  sample(fl, tx, ty, tex);
  for (int k = 0; k < 4; k++) out[k] = tex[k] * vc[k];
  for (int k = 0; k < 3; k++) out[k] *= vc[3];                      // colorTex.rgb *= color.a
}

__device__ void base3d_fog(const RasterFrame& rf, uint32_t inter, float px, float py, float out[4]) {
  const float dd = decode_depth(inter);
  float v = dd;
  bool curve_ok = true;
  if (rf.curve == T3D_FOG_CURVE_OUT_CIRC) v = sqrtf(1.0f - (dd - 1.0f) * (dd - 1.0f));
  else if (rf.curve == T3D_FOG_CURVE_IN_CIRC) v = 1.0f - sqrtf(1.0f - dd * dd);
  else if (rf.curve != T3D_FOG_CURVE_LINEAR) curve_ok = false;     // no curve matches: d stays 0
  float d = 0.0f;
  if (curve_ok) {                                                   // smoothstep(FogRange[0], FogRange[1], v)
    float t = (v - rf.fog_range[0]) / (rf.fog_range[1] - rf.fog_range[0]);
    t = t < 0.0f ? 0.0f : (t > 1.0f ? 1.0f : t);
    d = t * t * (3.0f - 2.0f * t);
  }
  if (rf.dither > 0) {                                              // step(0, d - BayerMatrix[...])
    const float bayer[16] = {1.0f / 17.0f,  9.0f / 17.0f, 3.0f / 17.0f,  11.0f / 17.0f, 13.0f / 17.0f, 5.0f / 17.0f,
                             15.0f / 17.0f, 7.0f / 17.0f, 4.0f / 17.0f,  12.0f / 17.0f, 2.0f / 17.0f,  10.0f / 17.0f,
                             16.0f / 17.0f, 8.0f / 17.0f, 14.0f / 17.0f, 6.0f / 17.0f};
    const int yc = ((int)(py / rf.dither)) % 4, xc = ((int)(px / rf.dither)) % 4;
    d = (d - bayer[yc * 4 + xc]) < 0.0f ? 0.0f : 1.0f;
  }
  const float al = out[3];
  if (rf.fog[3] == 0.0f) {
    for (int k = 0; k < 3; k++) out[k] += rf.fog[k] * d * al;
  } else if (rf.fog[3] == 1.0f) {
    for (int k = 0; k < 3; k++) out[k] -= rf.fog[k] * d * al;
  } else if (rf.fog[3] == 2.0f) {
    for (int k = 0; k < 3; k++) out[k] = mixf_(out[k], rf.fog[k], d) * al;
  }
}

// ---- fragments -------------------------------------------------------------------

// per drawn triangle: its record and the number of pixels of its box (its fragment slots)
__global__ __launch_bounds__(256) void k_rsetup(RArgs a) {
  const uint32_t t = blockIdx.x * 256 + threadIdx.x;
  if (t > a.ntri) return;
  uint64_t n = 0;
  if (t < a.ntri) {
    STri s;
    load_tri(a, t, s);
    if (setup(s, a.rf.width, a.rf.height)) n = (uint64_t)(s.bx1 - s.bx0 + 1) * (uint64_t)(s.by1 - s.by0 + 1);
    s.pad[0] = s.pad[1] = s.pad[2] = 0;
    const float4* src = reinterpret_cast<const float4*>(&s);
    float4* dst = reinterpret_cast<float4*>(a.recs + t);
    for (int k = 0; k < 10; k++) dst[k] = src[k];
  }
  a.counts[t] = n;   // counts[ntri] = 0 closes the scan
}

// The triangle owning fragment slot `slot` among [lo, hi] (offsets: exclusive
// scan of the box areas; triangles without pixels own no slot).
__device__ __forceinline__ uint32_t tri_of_slot(const RArgs& a, uint64_t slot, uint32_t lo, uint32_t hi) {
  if (hi < lo) hi = lo;
  hi += 1;   // largest t in [lo, hi) with offsets[t] <= slot
  while (hi - lo > 1) {
    const uint32_t mid = (lo + hi) / 2;
    if (a.offsets[mid] <= slot) lo = mid;
    else hi = mid;
  }
  return lo;
}

// a slot's pixel: boxes hold fewer than 2^31 pixels, so 32-bit division
__device__ __forceinline__ void slot_pixel(const STri& t, uint64_t local, int& x, int& y) {
  const uint32_t bw = (uint32_t)(t.bx1 - t.bx0 + 1), l = (uint32_t)local;
  const uint32_t row = l / bw;
  x = t.bx0 + (int)(l - row * bw);
  y = t.by0 + (int)row;
}

// one thread per 256 slots: the triangles of the first and last, so that the
// slot kernels' blocks start from them (thousands of independent searches
// instead of a search at the head of every block)
__global__ __launch_bounds__(256) void k_rblocks(RArgs a, uint32_t nblocks) {
  const uint32_t b = blockIdx.x * 256 + threadIdx.x;
  if (b >= nblocks) return;
  const uint64_t last = b * 256ull + 255 < a.nslots ? b * 256ull + 255 : a.nslots - 1;
  const uint64_t first = b * 256ull < last ? b * 256ull : last;   // both real slots, first <= last
  const uint32_t t0 = tri_of_slot(a, first, 0, a.ntri - 1);
  a.btri[b] = make_uint2(t0, tri_of_slot(a, last, t0, a.ntri - 1));
}

// A block's 256 slots: the triangles of its first and last slot bound the
// search of every thread's triangle, which runs over those triangles' offsets
// staged in LDS (or in memory when triangles without pixels stretch the
// range past 256).
__device__ __forceinline__ uint32_t block_tri(const RArgs& a, uint64_t slot) {
  __shared__ uint64_t s_off[256];
  const uint2 bt = a.btri[blockIdx.x];
  const uint32_t t0 = bt.x, t1 = bt.y;
  const uint64_t last = blockIdx.x * 256ull + 255 < a.nslots ? blockIdx.x * 256ull + 255 : a.nslots - 1;
  const uint64_t q = slot < a.nslots ? slot : last;
  if (t1 - t0 >= 256) return tri_of_slot(a, q, t0, t1);      // block-uniform
  s_off[threadIdx.x] = t0 + threadIdx.x <= t1 ? a.offsets[t0 + threadIdx.x] : ~0ull;
  __syncthreads();
  uint32_t lo = 0, hi = t1 - t0 + 1;                         // largest j in [lo, hi) with s_off[j] <= q
  while (hi - lo > 1) {
    const uint32_t mid = (lo + hi) / 2;
    if (s_off[mid] <= q) lo = mid;
    else hi = mid;
  }
  return t0 + lo;
}

// one thread per slot: is its pixel covered? Each wavefront's 64 answers
// are one mask word and its count.
__global__ __launch_bounds__(256) void k_rcover(RArgs a) {
  const uint64_t slot = blockIdx.x * 256ull + threadIdx.x;
  const uint32_t t = block_tri(a, slot);
  bool covered = false;
  if (slot < a.nslots) {
    const STri& r = a.recs[t];
    int x, y;
    slot_pixel(r, slot - a.offsets[t], x, y);
    covered = cover_test(r, x, y);
  }
  const uint64_t m = __ballot(covered);
  const uint32_t w = (uint32_t)(slot >> 6);   // < nwaves: the grid is nwaves / 4 blocks
  if ((threadIdx.x & 63) == 0) {
    a.masks[w] = m;
    a.wcount[w] = (uint32_t)__popcll(m);
  }
  if (slot == 0) a.wcount[a.nwaves] = 0;   // closes the scan
}

// each fragment, at its place from the scan: its pixel (the sort key) and its
// triangle and place in the triangle's box (the value)
__global__ __launch_bounds__(256) void k_rslots(RArgs a) {
  const uint64_t slot = blockIdx.x * 256ull + threadIdx.x;
  const uint32_t t = block_tri(a, slot);
  const uint32_t lane = threadIdx.x & 63;
  const uint64_t m = a.masks[slot >> 6];
  if (slot >= a.nslots || !((m >> lane) & 1)) return;
  const uint32_t i = a.wofs[slot >> 6] + (uint32_t)__popcll(m & ((1ull << lane) - 1));
  const uint64_t local = slot - a.offsets[t];
  int x, y;
  slot_pixel(a.recs[t], local, x, y);
  a.keys[i] = (uint32_t)y * (uint32_t)a.rf.width + (uint32_t)x;
  a.vals[i] = ((uint64_t)t << 32) | local;
}

// each fragment shaded, in pixel order: what the per-pixel resolve needs
__global__ __launch_bounds__(256) void k_rfrag(RArgs a) {
  const uint32_t i = blockIdx.x * 256 + threadIdx.x;
  if (i >= a.nfrags) return;
  const uint64_t v = a.svals[i];
  const STri& r = a.recs[(uint32_t)(v >> 32)];
  int x, y;
  slot_pixel(r, v & 0xFFFFFFFFull, x, y);
  double l[3];
  cover(r, x, y, l);
  const RasterFlush fl = a.flushes[r.flush];
  Frag fr;
  fr.flush = r.flush;
  fr.pad = 0;
  fr.d = interp(l, r.dz);
  fr.clip_a = 1.0f;
  if (a.rf.render_depth) {
    if (fl.flags & RF_CLIP) {                               // the clip-alpha depth shader (camera.go:441-478)
      float tx = interp(l, r.sx), ty = interp(l, r.sy);
      if (a.rf.pc_tex) {
        const float rc = 1.0f / interp(l, r.c0);
        tx *= rc;
        ty *= rc;
      }
      tx = fractf_(tx / (float)fl.tw) * (float)fl.tw;
      ty = fractf_(ty / (float)fl.th) * (float)fl.th;
      float tex[4];
      sample(fl, tx, ty, tex);
      fr.clip_a = tex[3] <= 0.8f ? -1.0f : tex[3];
    }
    const float vc[4] = {interp(l, r.r), interp(l, r.g), interp(l, r.b), interp(l, r.a)};
    base3d_color(fl, a.rf, vc, interp(l, r.sx), interp(l, r.sy), interp(l, r.c0), fr.col);
  } else {                                                  // DrawTriangles: texel times the premultiplied colour
    const float al = interp(l, r.a);
    const float vc[4] = {interp(l, r.r) * al, interp(l, r.g) * al, interp(l, r.b) * al, al};
    const float tx = modf_(interp(l, r.sx), (float)fl.tw), ty = modf_(interp(l, r.sy), (float)fl.th);
    float tex[4];
    sample(fl, tx, ty, tex);
    for (int q = 0; q < 4; q++) fr.col[q] = tex[q] * vc[q];
  }
  a.sfrags[i] = fr;
}

// runs of one pixel's fragments from one flush: where each starts ...
__global__ __launch_bounds__(256) void k_rseg_start(RArgs a) {
  const uint32_t k = blockIdx.x * 256 + threadIdx.x;
  if (k > a.nfrags) return;
  a.seg_start[k] = (k < a.nfrags && (k == 0 || a.skeys[k] != a.skeys[k - 1] ||
                                     a.sfrags[k].flush != a.sfrags[k - 1].flush)) ? 1u : 0u;
}
// ... and where it ends (seg_id[k] counts the starts before k: the run of fragment k is seg_id[k + 1] - 1)
__global__ __launch_bounds__(256) void k_rseg_end(RArgs a) {
  const uint32_t k = blockIdx.x * 256 + threadIdx.x;
  if (k >= a.nfrags) return;
  if (k + 1 == a.nfrags || a.seg_start[k + 1]) a.seg_end[a.seg_id[k + 1] - 1] = k + 1;
}

__global__ __launch_bounds__(256) void k_rranges(RArgs a) {
  const uint32_t k = blockIdx.x * 256 + threadIdx.x;
  if (k >= a.nfrags) return;
  const uint32_t key = a.skeys[k];
  if (k == 0 || a.skeys[k - 1] != key) a.pix_begin[key] = k;
  if (k + 1 == a.nfrags || a.skeys[k + 1] != key) a.pix_end[key] = k + 1;
}

// One thread per pixel: its fragments in draw order, flush by flush as flush
// draws them: the depth pass into the intermediate (tested against the
// result depth), the composite for opaque pairs, then the colour pass (fog
// from the intermediate) blended over the colour.  A blend of alpha 1
// replaces what is under it (src + dst * 0), so each pass starts at the last
// fragment that does so: the last passing depth fragment (the plain depth
// shader writes alpha 1) and the last colour fragment of alpha 1 — the same
// bits as blending every fragment in order.
__global__ __launch_bounds__(256) void k_resolve(RArgs a) {
  const uint32_t pix = blockIdx.x * 256 + threadIdx.x;
  const uint32_t W = (uint32_t)a.rf.width, H = (uint32_t)a.rf.height;
  if (pix >= W * H) return;
  const float px = (float)(pix % W) + 0.5f, py = (float)(pix / W) + 0.5f;   // dstPos
  const Frag* fs = a.sfrags;
  uint32_t col = a.img_color[pix];
  const uint32_t b = a.pix_begin[pix], e = a.pix_end[pix];
  if (!a.rf.render_depth) {
    uint32_t k0 = b;
    for (uint32_t k = e; k > b; k--)
      if (fs[k - 1].col[3] == 1.0f) {
        k0 = k - 1;
        break;
      }
    for (uint32_t k = k0; k < e; k++) col = source_over(col, fs[k].col);
    a.img_color[pix] = col;
    return;
  }
  uint32_t dep = a.img_depth[pix];
  for (uint32_t k = b, run = b < e ? a.seg_id[b + 1] - 1 : 0; k < e; run++) {
    const uint32_t f = fs[k].flush;
    const uint32_t k1 = a.seg_end[run];
    const RasterFlush fl = a.flushes[f];
    const bool empty_dep = ((dep >> 24) & 255u) == 0;
    const float dd = empty_dep ? 0.0f : decode_depth(dep);
    uint32_t inter = 0u;
    if (fl.flags & RF_CLIP) {                               // texel alphas below 1: every passing fragment blends
      for (uint32_t q = k; q < k1; q++) {
        if (fs[q].clip_a < 0.0f) continue;                  // discard(): texel alpha <= 0.8
        if (!(empty_dep || dd > fs[q].d)) continue;
        float enc[4];
        encode_depth(fs[q].d, enc);
        enc[3] = fs[q].clip_a;
        inter = source_over(inter, enc);
      }
    } else {
      for (uint32_t q = k1; q > k; q--)
        if (empty_dep || dd > fs[q - 1].d) {
          float enc[4];
          encode_depth(fs[q - 1].d, enc);
          inter = source_over(0u, enc);
          break;
        }
    }
    if (fl.flags & RF_COMPOSITE) {                          // resultDepthTexture.DrawImage(depthIntermediate)
      const float src[4] = {chan(inter, 0), chan(inter, 1), chan(inter, 2), chan(inter, 3)};
      dep = source_over(dep, src);
    }
    if (((inter >> 24) & 255u) != 0) {                     // colour pass; discard() where the intermediate is empty
      uint32_t j0 = k;
      for (uint32_t q = k1; q > k; q--)
        if (fs[q - 1].col[3] == 1.0f) {
          j0 = q - 1;
          break;
        }
      for (uint32_t q = j0; q < k1; q++) {
        float o[4] = {fs[q].col[0], fs[q].col[1], fs[q].col[2], fs[q].col[3]};
        if (!(fl.flags & RF_FOGLESS)) base3d_fog(a.rf, inter, px, py, o);
        col = source_over(col, o);
      }
    }
    k = k1;
  }
  a.img_color[pix] = col;
  a.img_depth[pix] = dep;
}

uint32_t host_q8(float x) {   // the device's q8, on the host for the clear words
  if (!(x > 0.0f)) x = 0.0f;
  if (x > 1.0f) x = 1.0f;
  return (uint32_t)std::floor(x * 255.0f + 0.5f);
}

}  // namespace
}  // namespace t3d

using namespace t3d;

#define R_TRY(expr)                                                                     \
  do {                                                                                  \
    hipError_t e_ = (expr);                                                             \
    if (e_ != hipSuccess) return fail(c, T3D_ERR_DEVICE, "%s: %s", #expr, hipGetErrorString(e_)); \
  } while (0)

extern "C" {

t3d_status t3d_texture_upload(t3d_ctx* c, uint32_t width, uint32_t height, const uint8_t* rgba8, uint64_t* out) {
  if (!c || !rgba8 || !out || !width || !height) return T3D_ERR_INVALID_ARGUMENT;
  std::lock_guard<std::mutex> lk(c->mu);
  (void)hipSetDevice(c->device);
  auto tex = std::make_unique<Texture>();
  const uint64_t bytes = 4ull * width * height;
  t3d_status st = dev_alloc(c, tex->buf, bytes);
  if (st) return st;
  R_TRY(hipMemcpy(tex->buf.ptr, rgba8, bytes, hipMemcpyHostToDevice));
  tex->w = width;
  tex->h = height;
  tex->handle = c->next_handle++;
  *out = tex->handle;
  c->textures[tex->handle] = std::move(tex);
  return T3D_OK;
}

t3d_status t3d_texture_free(t3d_ctx* c, uint64_t handle) {
  if (!c) return T3D_ERR_INVALID_ARGUMENT;
  std::lock_guard<std::mutex> lk(c->mu);
  auto it = c->textures.find(handle);
  if (it == c->textures.end()) return fail(c, T3D_ERR_NOT_FOUND, "texture_free: unknown handle");
  (void)hipDeviceSynchronize();   // a raster call in flight may still read it
  dev_free(c, it->second->buf);
  c->textures.erase(it);
  return T3D_OK;
}

t3d_status t3d_frame_rasterize(t3d_ctx* c, t3d_frame* f) {
  if (!c || !f) return T3D_ERR_INVALID_ARGUMENT;
  std::lock_guard<std::mutex> lk(c->mu);
  (void)hipSetDevice(c->device);
  if (f->inflight) return fail(c, T3D_ERR_NOT_READY, "rasterize: the frame is in flight (t3d_frame_wait first)");
  if (f->fcs.size() != 1) return fail(c, T3D_ERR_UNSUPPORTED, "rasterize: a frame of one Camera.Render only");
  const RasterFrame rf = f->rframe;
  const auto& flushes = f->out_pairs();
  if (flushes.size() != f->rflush.size()) return fail(c, T3D_ERR_INVALID_ARGUMENT, "rasterize: no rendered frame");
  for (size_t g = 0; g < flushes.size(); g++) {
    if (!flushes[g].survivor_count) continue;
    if (f->rflush[g].flags & RF_BAD_BLEND)
      return fail(c, T3D_ERR_UNSUPPORTED, "rasterize: flush %zu blends other than SourceOver", g);
    if (f->rflush[g].flags & RF_NO_TEXTURE)
      return fail(c, T3D_ERR_NOT_FOUND, "rasterize: flush %zu names an unknown texture", g);
  }
  const uint32_t W = (uint32_t)rf.width, H = (uint32_t)rf.height;
  // textures are resolved again: one freed since the render is refused rather than read
  std::vector<RasterFlush> rfl = f->rflush;
  for (size_t g = 0; g < rfl.size(); g++) {
    if (!f->rflush_tex[g] || !flushes[g].survivor_count) continue;
    auto it = c->textures.find(f->rflush_tex[g]);
    if (it == c->textures.end())
      return fail(c, T3D_ERR_NOT_FOUND, "rasterize: flush %zu names an unknown texture", g);
    rfl[g].tex = reinterpret_cast<const uint8_t*>(it->second->buf.ptr);
    rfl[g].tw = it->second->w;
    rfl[g].th = it->second->h;
  }
  // render pairs with survivors, their flush and their first drawn triangle
  std::vector<RasterPair> rp;
  {
    uint32_t tri = 0, g = 0;
    for (uint32_t i = 0; i < f->nparts; i++) {
      if (f->grouped)
        while (g + 1 < f->flush_begin.size() && f->flush_begin[g + 1] <= i) g++;
      const t3d_pair_result& pr = f->pairs[i];
      if (!pr.survivor_count) continue;
      rp.push_back(RasterPair{tri, f->grouped ? g : i, pr.vertex_offset});
      tri += pr.survivor_count;
    }
  }
  const uint64_t ntri64 = f->total_survivors;
  if (ntri64 >= 0xFFFFFFFFull) return fail(c, T3D_ERR_UNSUPPORTED, "rasterize: too many triangles");
  const uint32_t ntri = (uint32_t)ntri64;
  hipStream_t st = c->stream;
  // images
  const uint64_t img = 4ull * W * H;
  t3d_status s = ensure(c, f->images, f->images_cap, img * (rf.render_depth ? 2 : 1));
  if (s) return s;
  f->image_color = reinterpret_cast<uint8_t*>(f->images.ptr);
  f->image_depth = rf.render_depth ? f->image_color + img : nullptr;
  f->image_w = rf.width;
  f->image_h = rf.height;
  const float* cl = rf.clear;
  const uint32_t clear_word = host_q8(cl[0] * cl[3]) | (host_q8(cl[1] * cl[3]) << 8) | (host_q8(cl[2] * cl[3]) << 16) |
                              (host_q8(cl[3]) << 24);
  R_TRY(launch_fill(f->image_color, img, clear_word, st));
  if (f->image_depth) R_TRY(launch_fill(f->image_depth, img, 0u, st));
  if (ntri == 0) {
    R_TRY(hipStreamSynchronize(st));
    return T3D_OK;
  }
  // phase 1: tables, triangle records, box areas and their scan
  size_t scan_tmp = 0;
  R_TRY(hipcub::DeviceScan::ExclusiveSum(nullptr, scan_tmp, (const uint64_t*)nullptr, (uint64_t*)nullptr, ntri + 1, st));
  auto al = [](uint64_t x) { return (x + 255) / 256 * 256; };
  const uint64_t b_pairs = al(sizeof(RasterPair) * rp.size()), b_fl = al(sizeof(RasterFlush) * rfl.size());
  const uint64_t b_cnt = al(8ull * (ntri + 1)), b_recs = al(sizeof(STri) * (uint64_t)ntri);
  s = ensure(c, f->raster_buf, f->raster_cap, b_pairs + b_fl + 2 * b_cnt + b_recs + al(scan_tmp));
  if (s) return s;
  uint8_t* w = reinterpret_cast<uint8_t*>(f->raster_buf.ptr);
  RArgs a{};
  a.pairs = reinterpret_cast<RasterPair*>(w);
  a.npairs = (uint32_t)rp.size();
  a.ntri = ntri;
  a.flushes = reinterpret_cast<RasterFlush*>(w + b_pairs);
  a.color = f->args.out_color;
  a.depth = f->args.out_depth;
  a.rf = rf;
  a.counts = reinterpret_cast<uint64_t*>(w + b_pairs + b_fl);
  uint64_t* offsets = reinterpret_cast<uint64_t*>(w + b_pairs + b_fl + b_cnt);
  a.offsets = offsets;
  a.recs = reinterpret_cast<STri*>(w + b_pairs + b_fl + 2 * b_cnt);
  void* scan_mem = w + b_pairs + b_fl + 2 * b_cnt + b_recs;
  a.img_color = reinterpret_cast<uint32_t*>(f->image_color);
  a.img_depth = reinterpret_cast<uint32_t*>(f->image_depth);
  R_TRY(hipMemcpyAsync((void*)a.pairs, rp.data(), sizeof(RasterPair) * rp.size(), hipMemcpyHostToDevice, st));
  R_TRY(hipMemcpyAsync((void*)a.flushes, rfl.data(), sizeof(RasterFlush) * rfl.size(), hipMemcpyHostToDevice, st));
  const uint32_t tblocks = (ntri + 1 + 255) / 256;
  hipLaunchKernelGGL(k_rsetup, dim3(tblocks), dim3(256), 0, st, a);
  R_TRY(hipGetLastError());
  R_TRY(hipcub::DeviceScan::ExclusiveSum(scan_mem, scan_tmp, a.counts, offsets, ntri + 1, st));
  uint64_t nslots = 0;
  R_TRY(hipMemcpyAsync(&nslots, offsets + ntri, 8, hipMemcpyDeviceToHost, st));
  R_TRY(hipStreamSynchronize(st));
  if (nslots >= 0x7FFFFFFFull)
    return fail(c, T3D_ERR_UNSUPPORTED, "rasterize: %llu box pixels (at most 2^31 - 1)", (unsigned long long)nslots);
  a.nslots = nslots;
  if (nslots) {
    // phase 2: coverage masks, 64 slots a word, and the scan of their counts: each covered slot's fragment
    const uint32_t sblocks = (uint32_t)((nslots + 255) / 256);
    a.nwaves = 4 * sblocks;
    size_t fscan_tmp = 0;
    R_TRY(hipcub::DeviceScan::ExclusiveSum(nullptr, fscan_tmp, (const uint32_t*)nullptr, (uint32_t*)nullptr,
                                           (int)(a.nwaves + 1), st));
    const uint64_t b_masks = al(8ull * a.nwaves), b_wc = al(4ull * (a.nwaves + 1));
    s = ensure(c, f->raster_buf2, f->raster_cap2, b_masks + 2 * b_wc + al(fscan_tmp) + al(8ull * sblocks));
    if (s) return s;
    uint8_t* w2 = reinterpret_cast<uint8_t*>(f->raster_buf2.ptr);
    a.masks = reinterpret_cast<uint64_t*>(w2);
    a.wcount = reinterpret_cast<uint32_t*>(w2 + b_masks);
    uint32_t* wofs = reinterpret_cast<uint32_t*>(w2 + b_masks + b_wc);
    a.wofs = wofs;
    a.btri = reinterpret_cast<uint2*>(w2 + b_masks + 2 * b_wc + al(fscan_tmp));
    hipLaunchKernelGGL(k_rblocks, dim3((sblocks + 255) / 256), dim3(256), 0, st, a, sblocks);
    R_TRY(hipGetLastError());
    hipLaunchKernelGGL(k_rcover, dim3(sblocks), dim3(256), 0, st, a);
    R_TRY(hipGetLastError());
    R_TRY(hipcub::DeviceScan::ExclusiveSum(w2 + b_masks + 2 * b_wc, fscan_tmp, a.wcount, wofs, (int)(a.nwaves + 1),
                                           st));
    uint32_t nfrags = 0;
    R_TRY(hipMemcpyAsync(&nfrags, wofs + a.nwaves, 4, hipMemcpyDeviceToHost, st));
    R_TRY(hipStreamSynchronize(st));
    a.nfrags = nfrags;
    if (nfrags > 0) {
      // phase 3: shading, the sort by pixel, the fragments in that order, each pixel's range, the resolve
      uint32_t bits = 1;
      while ((1ull << bits) < (uint64_t)W * H) bits++;
      size_t sort_tmp = 0;
      R_TRY(hipcub::DeviceRadixSort::SortPairs(nullptr, sort_tmp, (const uint32_t*)nullptr, (uint32_t*)nullptr,
                                               (const uint64_t*)nullptr, (uint64_t*)nullptr, (int)nfrags, 0, (int)bits,
                                               st));
      const uint64_t b_fr = al(sizeof(Frag) * (uint64_t)nfrags), b_k = al(4ull * nfrags), b_pix = al(4ull * W * H);
      size_t seg_tmp = 0;
      R_TRY(hipcub::DeviceScan::ExclusiveSum(nullptr, seg_tmp, (const uint32_t*)nullptr, (uint32_t*)nullptr,
                                             (int)nfrags + 1, st));
      const uint64_t b_k1 = al(4ull * (nfrags + 1));
      const uint64_t tail = std::max<uint64_t>(al(sort_tmp), al(seg_tmp));   // the two passes share it
      s = ensure(c, f->raster_buf3, f->raster_cap3, b_fr + 6 * b_k + 2 * b_pix + 3 * b_k1 + tail);
      if (s) return s;
      uint8_t* w3 = reinterpret_cast<uint8_t*>(f->raster_buf3.ptr);
      uint8_t* q = w3;
      auto take = [&](uint64_t n) { uint8_t* r = q; q += n; return r; };
      a.sfrags = reinterpret_cast<Frag*>(take(b_fr));
      a.keys = reinterpret_cast<uint32_t*>(take(b_k));
      uint32_t* sk = reinterpret_cast<uint32_t*>(take(b_k));
      a.vals = reinterpret_cast<uint64_t*>(take(2 * b_k));
      uint64_t* sv = reinterpret_cast<uint64_t*>(take(2 * b_k));
      a.skeys = sk;
      a.svals = sv;
      a.pix_begin = reinterpret_cast<uint32_t*>(take(b_pix));
      a.pix_end = reinterpret_cast<uint32_t*>(take(b_pix));
      a.seg_start = reinterpret_cast<uint32_t*>(take(b_k1));
      uint32_t* seg_id = reinterpret_cast<uint32_t*>(take(b_k1));
      a.seg_id = seg_id;
      a.seg_end = reinterpret_cast<uint32_t*>(take(b_k1));
      uint8_t* tmp3 = take(tail);
      const uint32_t fblocks = (nfrags + 255) / 256;
      hipLaunchKernelGGL(k_rslots, dim3(sblocks), dim3(256), 0, st, a);
      R_TRY(hipGetLastError());
      R_TRY(hipcub::DeviceRadixSort::SortPairs(tmp3, sort_tmp, a.keys, sk, a.vals, sv, (int)nfrags, 0, (int)bits, st));
      hipLaunchKernelGGL(k_rfrag, dim3(fblocks), dim3(256), 0, st, a);
      R_TRY(hipGetLastError());
      hipLaunchKernelGGL(k_rseg_start, dim3((nfrags + 1 + 255) / 256), dim3(256), 0, st, a);
      R_TRY(hipGetLastError());
      R_TRY(hipcub::DeviceScan::ExclusiveSum(tmp3, seg_tmp, a.seg_start, seg_id, (int)nfrags + 1, st));
      hipLaunchKernelGGL(k_rseg_end, dim3(fblocks), dim3(256), 0, st, a);
      R_TRY(hipGetLastError());
      R_TRY(launch_fill(a.pix_begin, 2 * b_pix, 0u, st));   // pix_begin and pix_end: no fragments
      hipLaunchKernelGGL(k_rranges, dim3(fblocks), dim3(256), 0, st, a);
      R_TRY(hipGetLastError());
      hipLaunchKernelGGL(k_resolve, dim3((W * H + 255) / 256), dim3(256), 0, st, a);
      R_TRY(hipGetLastError());
    }
  }
  R_TRY(hipStreamSynchronize(st));
  return T3D_OK;
}

t3d_status t3d_frame_device_images(const t3d_frame* f, const uint8_t** color, const uint8_t** depth) {
  if (!f) return T3D_ERR_INVALID_ARGUMENT;
  if (!f->image_color) return T3D_ERR_NOT_FOUND;
  if (color) *color = f->image_color;
  if (depth) *depth = f->image_depth;
  return T3D_OK;
}

t3d_status t3d_frame_download_images(t3d_ctx* c, const t3d_frame* f, uint8_t* color, uint8_t* depth) {
  if (!c || !f) return T3D_ERR_INVALID_ARGUMENT;
  std::lock_guard<std::mutex> lk(c->mu);
  (void)hipSetDevice(c->device);
  if (!f->image_color) return fail(c, T3D_ERR_NOT_FOUND, "download_images: the frame was not rasterized");
  if (depth && !f->image_depth) return fail(c, T3D_ERR_NOT_FOUND, "download_images: no depth texture (RenderDepth off)");
  const uint64_t img = 4ull * f->image_w * f->image_h;
  if (color) R_TRY(hipMemcpy(color, f->image_color, img, hipMemcpyDeviceToHost));
  if (depth) R_TRY(hipMemcpy(depth, f->image_depth, img, hipMemcpyDeviceToHost));
  return T3D_OK;
}

}  // extern "C"
