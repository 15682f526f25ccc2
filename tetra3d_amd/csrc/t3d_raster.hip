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
// Work split: every drawn triangle is binned by its pixel bounding box into
// 16 x 16 tiles (count, exclusive scan, emit, stable radix sort by tile), so a
// tile's list is in draw order.  One 256-thread block per tile then walks its
// list, one thread per pixel, the pixel's colour, depth and depth-intermediate
// values held in registers from the clear to the final store: per flush
// segment a depth pass, the intermediate composited over the depth (opaque
// pairs), then the colour pass through base3d.
#include <hip/hip_runtime.h>
#include <hipcub/hipcub.hpp>

#include <cmath>
#include <cstring>

#include "t3d_internal.h"
#include "t3d_runtime.h"

namespace t3d {
namespace {

constexpr int kTileEdge = 16;     // pixels per tile edge: one 256-thread block per tile
constexpr uint32_t kStage = 64;   // triangles staged in LDS per round

struct RArgs {
  const RasterPair* pairs;        // render pairs with survivors, in draw order
  uint32_t npairs;
  uint32_t ntri;                  // drawn triangles of the frame
  const RasterFlush* flushes;
  const float4* color;            // frame outputs, t3d_vertex as 3 float4
  const float4* depth;
  RasterFrame rf;
  uint32_t tiles_x, tiles_y;
  uint64_t* counts;               // [ntri + 1] tiles each triangle's box touches
  const uint64_t* offsets;        // [ntri + 1] exclusive scan of counts
  uint32_t* keys;                 // [total] tile of each (triangle, tile) entry
  uint32_t* vals;                 // [total] its triangle
  const uint32_t* skeys;          // the entries sorted by tile, stable
  const uint32_t* svals;
  uint32_t* tile_begin;           // [tiles] range of each tile in the sorted entries
  uint32_t* tile_end;
  uint32_t total;
  uint32_t* img_color;            // RGBA8 packed, r in the low byte
  uint32_t* img_depth;
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

// A triangle as staged: corner positions and the attributes the draws read.
struct STri {
  float x[3], y[3];          // DstX, DstY
  float sx[3], sy[3], c0[3]; // SrcX, SrcY, Custom0 (the same on both lists)
  float r[3], g[3], b[3], a[3];
  float dz[3];               // the depth list's ColorR
  double area, s;
  int bx0, bx1, by0, by1;    // pixel box on the image; bx0 > bx1: nothing to draw
  uint32_t flush;
};

// Area, orientation and the box of pixels whose centres the corners span.
__device__ __forceinline__ bool setup(STri& t, int W, int H) {
  t.area = edge_fn(t.x[0], t.y[0], t.x[1], t.y[1], (double)t.x[2], (double)t.y[2]);
  t.bx0 = 1;
  t.bx1 = 0;
  if (!(t.area != 0.0) || t.area != t.area) return false;
  t.s = t.area > 0 ? 1.0 : -1.0;
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

__device__ __forceinline__ bool cover(const STri& t, int x, int y, double l[3]) {
  if (x < t.bx0 || x > t.bx1 || y < t.by0 || y > t.by1) return false;
  const double cx = x + 0.5, cy = y + 0.5, s = t.s;
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

// ---- binning --------------------------------------------------------------------

__global__ __launch_bounds__(256) void k_rcount(RArgs a) {
  const uint32_t t = blockIdx.x * 256 + threadIdx.x;
  if (t > a.ntri) return;
  uint64_t n = 0;
  if (t < a.ntri) {
    STri s;
    load_tri(a, t, s);
    if (setup(s, a.rf.width, a.rf.height))
      n = (uint64_t)(s.bx1 / kTileEdge - s.bx0 / kTileEdge + 1) * (uint64_t)(s.by1 / kTileEdge - s.by0 / kTileEdge + 1);
  }
  a.counts[t] = n;   // counts[ntri] = 0 closes the scan
}

__global__ __launch_bounds__(256) void k_remit(RArgs a) {
  const uint32_t t = blockIdx.x * 256 + threadIdx.x;
  if (t >= a.ntri || a.counts[t] == 0) return;
  STri s;
  load_tri(a, t, s);
  if (!setup(s, a.rf.width, a.rf.height)) return;
  uint64_t o = a.offsets[t];
  for (int ty = s.by0 / kTileEdge; ty <= s.by1 / kTileEdge; ty++)
    for (int tx = s.bx0 / kTileEdge; tx <= s.bx1 / kTileEdge; tx++) {
      a.keys[o] = (uint32_t)ty * a.tiles_x + (uint32_t)tx;
      a.vals[o] = t;
      o++;
    }
}

__global__ __launch_bounds__(256) void k_rranges(RArgs a) {
  const uint32_t k = blockIdx.x * 256 + threadIdx.x;
  if (k >= a.total) return;
  const uint32_t key = a.skeys[k];
  if (k == 0 || a.skeys[k - 1] != key) a.tile_begin[key] = k;
  if (k + 1 == a.total || a.skeys[k + 1] != key) a.tile_end[key] = k + 1;
}

// ---- drawing ---------------------------------------------------------------------

// base3d.kage Fragment (61-153) at a pixel whose depth intermediate has alpha > 0
__device__ void base3d(const RasterFlush& fl, const RasterFrame& rf, bool fogless, uint32_t inter, const float vc[4],
                       float sx, float sy, float c0, float px, float py, float out[4]) {
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
  if (fogless) return;
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

// Stages the sorted entries [p, p + n) of this tile; returns, through s_lead,
// how many of them from the first belong to flush `f` (f == ~0u: the first's).
__device__ void stage(const RArgs& a, uint32_t p, uint32_t n, STri* s_tri, uint32_t* s_lead) {
  const int W = a.rf.width, H = a.rf.height;
  if (threadIdx.x < n) {
    STri& s = s_tri[threadIdx.x];
    load_tri(a, a.svals[p + threadIdx.x], s);
    setup(s, W, H);
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    uint32_t m = 0;
    while (m < n && s_tri[m].flush == s_tri[0].flush) m++;
    *s_lead = m;
  }
  __syncthreads();
}

__global__ __launch_bounds__(256) void k_raster(RArgs a) {
  __shared__ STri s_tri[kStage];
  __shared__ uint32_t s_lead;
  const uint32_t tile = blockIdx.x;
  const int W = a.rf.width, H = a.rf.height;
  const int x = (int)(tile % a.tiles_x) * kTileEdge + (int)(threadIdx.x % kTileEdge);
  const int y = (int)(tile / a.tiles_x) * kTileEdge + (int)(threadIdx.x / kTileEdge);
  const bool valid = x < W && y < H;
  const size_t pix = valid ? (size_t)y * W + x : 0;
  uint32_t col = valid ? a.img_color[pix] : 0u;
  uint32_t dep = (valid && a.rf.render_depth) ? a.img_depth[pix] : 0u;
  const float fx = (float)x + 0.5f, fy = (float)y + 0.5f;   // dstPos
  const uint32_t b = a.tile_begin[tile], e = a.tile_end[tile];
  uint32_t pos = b;
  while (pos < e) {                                         // block-uniform: one flush segment per round
    uint32_t n = min(kStage, e - pos);
    stage(a, pos, n, s_tri, &s_lead);
    const uint32_t f = s_tri[0].flush;
    const RasterFlush fl = a.flushes[f];
    if (!a.rf.render_depth) {
      // DrawTriangles(colorVertexList, img): texel times the premultiplied vertex colour
      const uint32_t m = s_lead;
      for (uint32_t i = 0; i < m && valid; i++) {
        const STri& t = s_tri[i];
        double l[3];
        if (!cover(t, x, y, l)) continue;
        const float al = interp(l, t.a);
        const float vc[4] = {interp(l, t.r) * al, interp(l, t.g) * al, interp(l, t.b) * al, al};
        const float tx = modf_(interp(l, t.sx), (float)fl.tw), ty = modf_(interp(l, t.sy), (float)fl.th);
        float tex[4], o[4];
        sample(fl, tx, ty, tex);
        for (int q = 0; q < 4; q++) o[q] = tex[q] * vc[q];
        col = source_over(col, o);
      }
      __syncthreads();
      pos += m;
      continue;
    }
    // depth pass over the segment, chunk by chunk, into this pixel's intermediate
    uint32_t inter = 0u;
    const uint32_t seg0 = pos;
    uint32_t p = pos, m = s_lead;
    for (;;) {
      for (uint32_t i = 0; i < m && valid; i++) {
        const STri& t = s_tri[i];
        double l[3];
        if (!cover(t, x, y, l)) continue;
        const float d = interp(l, t.dz);
        if (!(((dep >> 24) & 255u) == 0 || decode_depth(dep) > d)) continue;
        float enc[4];
        encode_depth(d, enc);
        if (fl.flags & RF_CLIP) {                           // the clip-alpha depth shader (camera.go:441-478)
          float tx = interp(l, t.sx), ty = interp(l, t.sy);
          if (a.rf.pc_tex) {
            const float r = 1.0f / interp(l, t.c0);
            tx *= r;
            ty *= r;
          }
          tx = fractf_(tx / (float)fl.tw) * (float)fl.tw;
          ty = fractf_(ty / (float)fl.th) * (float)fl.th;
          float tex[4];
          sample(fl, tx, ty, tex);
          if (tex[3] <= 0.8f) continue;
          enc[3] = tex[3];
        }
        inter = source_over(inter, enc);
      }
      __syncthreads();
      p += m;
      if (m < n || p >= e) break;                           // the segment ends in this chunk
      n = min(kStage, e - p);
      stage(a, p, n, s_tri, &s_lead);
      m = s_tri[0].flush == f ? s_lead : 0u;
      if (m == 0) break;
    }
    const uint32_t seg1 = p;
    if (fl.flags & RF_COMPOSITE) {                          // resultDepthTexture.DrawImage(depthIntermediate)
      const float src[4] = {chan(inter, 0), chan(inter, 1), chan(inter, 2), chan(inter, 3)};
      dep = source_over(dep, src);
    }
    // colour pass through base3d over the same segment
    const bool fogless = (fl.flags & RF_FOGLESS) != 0;
    for (p = seg0; p < seg1;) {
      n = min(kStage, seg1 - p);
      stage(a, p, n, s_tri, &s_lead);
      for (uint32_t i = 0; i < n && valid; i++) {
        const STri& t = s_tri[i];
        double l[3];
        if (((inter >> 24) & 255u) == 0 || !cover(t, x, y, l)) continue;   // discard()
        const float vc[4] = {interp(l, t.r), interp(l, t.g), interp(l, t.b), interp(l, t.a)};
        float o[4];
        base3d(fl, a.rf, fogless, inter, vc, interp(l, t.sx), interp(l, t.sy), interp(l, t.c0), fx, fy, o);
        col = source_over(col, o);
      }
      __syncthreads();
      p += n;
    }
    pos = seg1;
  }
  if (valid) {
    a.img_color[pix] = col;
    if (a.rf.render_depth) a.img_depth[pix] = dep;
  }
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
  const uint32_t tiles_x = (W + kTileEdge - 1) / kTileEdge, tiles_y = (H + kTileEdge - 1) / kTileEdge;
  const uint32_t ntiles = tiles_x * tiles_y;
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
  // working memory, phase 1: tables, counts, offsets, scan storage
  size_t scan_tmp = 0;
  R_TRY(hipcub::DeviceScan::ExclusiveSum(nullptr, scan_tmp, (const uint64_t*)nullptr, (uint64_t*)nullptr, ntri + 1, st));
  auto al = [](uint64_t x) { return (x + 255) / 256 * 256; };
  const uint64_t b_pairs = al(sizeof(RasterPair) * rp.size()), b_fl = al(sizeof(RasterFlush) * rfl.size());
  const uint64_t b_cnt = al(8ull * (ntri + 1)), b_tiles = al(4ull * ntiles);
  const uint64_t phase1 = b_pairs + b_fl + 2 * b_cnt + 2 * b_tiles + al(scan_tmp);
  s = ensure(c, f->raster_buf, f->raster_cap, phase1);
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
  a.tiles_x = tiles_x;
  a.tiles_y = tiles_y;
  a.counts = reinterpret_cast<uint64_t*>(w + b_pairs + b_fl);
  uint64_t* offsets = reinterpret_cast<uint64_t*>(w + b_pairs + b_fl + b_cnt);
  a.offsets = offsets;
  a.tile_begin = reinterpret_cast<uint32_t*>(w + b_pairs + b_fl + 2 * b_cnt);
  a.tile_end = reinterpret_cast<uint32_t*>(w + b_pairs + b_fl + 2 * b_cnt + b_tiles);
  void* scan_mem = w + b_pairs + b_fl + 2 * b_cnt + 2 * b_tiles;
  a.img_color = reinterpret_cast<uint32_t*>(f->image_color);
  a.img_depth = reinterpret_cast<uint32_t*>(f->image_depth);
  R_TRY(hipMemcpyAsync((void*)a.pairs, rp.data(), sizeof(RasterPair) * rp.size(), hipMemcpyHostToDevice, st));
  R_TRY(hipMemcpyAsync((void*)a.flushes, rfl.data(), sizeof(RasterFlush) * rfl.size(), hipMemcpyHostToDevice, st));
  R_TRY(launch_fill(a.tile_begin, 2 * b_tiles, 0u, st));   // tile_begin and tile_end: empty tiles
  const uint32_t tblocks = (ntri + 1 + 255) / 256;
  hipLaunchKernelGGL(k_rcount, dim3(tblocks), dim3(256), 0, st, a);
  R_TRY(hipGetLastError());
  R_TRY(hipcub::DeviceScan::ExclusiveSum(scan_mem, scan_tmp, a.counts, offsets, ntri + 1, st));
  uint64_t total64 = 0;
  R_TRY(hipMemcpyAsync(&total64, offsets + ntri, 8, hipMemcpyDeviceToHost, st));
  R_TRY(hipStreamSynchronize(st));
  if (total64 >= 0x80000000ull) return fail(c, T3D_ERR_UNSUPPORTED, "rasterize: %llu tile entries", (unsigned long long)total64);
  const uint32_t total = (uint32_t)total64;
  a.total = total;
  if (total) {
    // phase 2: entries and their sort
    uint32_t bits = 1;
    while ((1u << bits) < ntiles) bits++;
    size_t sort_tmp = 0;
    R_TRY(hipcub::DeviceRadixSort::SortPairs(nullptr, sort_tmp, (const uint32_t*)nullptr, (uint32_t*)nullptr,
                                             (const uint32_t*)nullptr, (uint32_t*)nullptr, total, 0, (int)bits, st));
    const uint64_t b_e = al(4ull * total);
    s = ensure(c, f->raster_buf2, f->raster_cap2, 4 * b_e + al(sort_tmp));
    if (s) return s;
    uint8_t* w2 = reinterpret_cast<uint8_t*>(f->raster_buf2.ptr);
    a.keys = reinterpret_cast<uint32_t*>(w2);
    a.vals = reinterpret_cast<uint32_t*>(w2 + b_e);
    uint32_t* sk = reinterpret_cast<uint32_t*>(w2 + 2 * b_e);
    uint32_t* sv = reinterpret_cast<uint32_t*>(w2 + 3 * b_e);
    a.skeys = sk;
    a.svals = sv;
    hipLaunchKernelGGL(k_remit, dim3(tblocks), dim3(256), 0, st, a);
    R_TRY(hipGetLastError());
    R_TRY(hipcub::DeviceRadixSort::SortPairs(w2 + 4 * b_e, sort_tmp, a.keys, sk, a.vals, sv, total, 0, (int)bits, st));
    hipLaunchKernelGGL(k_rranges, dim3((total + 255) / 256), dim3(256), 0, st, a);
    R_TRY(hipGetLastError());
  }
  hipLaunchKernelGGL(k_raster, dim3(ntiles), dim3(256), 0, st, a);
  R_TRY(hipGetLastError());
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
