// t3d_kernels.hip — CDNA4 (gfx950) kernels of the per-frame geometry path.
//
// One frame is a chain of launches over all render pairs ("parts") at once;
// every block works on exactly one part, found by binary search over a
// per-part prefix of block counts, so per-part uniforms are wave-uniform
// (scalar loads) and no block straddles two parts.
//
//   k_vertex    per vertex: shape keys, skinning or MVP transform, snapping,
//               lighting (lights staged in LDS), vertex colour, fog
//               (model.go:751-822, camera.go:1610-1644, 1740-1810)
//   k_cull      per triangle: w clamp, z-range / screen / backface rejects,
//               depth key, per-part survivor count and min/max depth
//               (model.go:828-969)
//   k_groups    dynamic batches: each later part of one flush placed after
//               the survivors of the parts before it (camera.go:2075-2099)
//   k_hist      per tile of a multi-tile part: 512-bin histogram
//   k_scan*     exclusive scan of each part's [bin][tile] histogram matrix
//   k_scatter   stable counting sort: output slot of every survivor
//               (sortingTriangleBucket Sort + ForEach, sortingTriangles.go:38-125)
//   k_assemble  per drawn triangle in slot order: ebiten colour/depth vertices
//               and indices, staged in LDS and written as contiguous float4 rows
//               with streaming stores (camera.go:1646-1820)
//   k_normals   Camera.RenderNormals: the normal vertex list (camera.go:1710-1722)
//
// All arithmetic is float32 in the reference's order; the build uses
// -ffp-contract=off and correctly rounded division and square root.
#include <hip/hip_runtime.h>

#include "t3d_internal.h"
#include "t3d_stage.h"
#include "t3d_anim.h"

namespace t3d {

constexpr float kMaxF32 = 3.40282346638528859811704183484516925440e+38f;

__device__ __forceinline__ uint32_t find_part(const uint32_t* __restrict__ prefix, uint32_t n, uint32_t b) {
  // Largest p in [0, n) with prefix[p] <= b (parts with zero blocks are skipped).
  uint32_t lo = 0, hi = n;
  while (hi - lo > 1) {
    uint32_t mid = (lo + hi) >> 1;
    if (prefix[mid] <= b) lo = mid;
    else hi = mid;
  }
  return lo;
}

// Workgroups of a launch are dealt round-robin to the 8 XCDs (workgroup i to
// XCD i % 8), each with its own L2.  With the swizzle on, the workgroups an
// XCD receives take consecutive logical block numbers, so the blocks of one
// render pair (which gather the same vertex records) share an L2.  The map
// is a bijection of [0, n).
constexpr uint32_t kXcds = 8;
__device__ __forceinline__ uint32_t block_index_n(uint32_t swizzle, uint32_t n) {
  const uint32_t i = blockIdx.x;
  if (!swizzle) return i;
  const uint32_t per = n / kXcds, rem = n % kXcds, x = i % kXcds;
  return x * per + min(x, rem) + i / kXcds;
}
__device__ __forceinline__ uint32_t block_index(uint32_t swizzle) { return block_index_n(swizzle, gridDim.x); }

// The logical block of this workgroup.  A frame built on the device is
// launched for every candidate pair being drawn; its actual block count is
// FrameCounts word `which` (times `mult`), and a workgroup past it gets
// kNoBlock (block-uniform: the whole workgroup returns).  The swizzle then
// deals the actual blocks, not the launch's, over the XCDs.
constexpr uint32_t kNoBlock = 0xFFFFFFFFu;
__device__ __forceinline__ uint32_t frame_block(const FrameArgs& a, uint32_t which, uint32_t mult) {
  uint32_t n = gridDim.x;
  if (a.counts) {
    n = a.counts[which] * mult;
    if (blockIdx.x >= n) return kNoBlock;
  }
  return block_index_n(a.xcd_swizzle, n);
}
// The same for the launches that never swizzle (scan blocks, pairs).
__device__ __forceinline__ bool past_count(const FrameArgs& a, uint32_t which) {
  return a.counts && blockIdx.x >= a.counts[which];
}

__device__ __forceinline__ float clampf(float v, float lo, float hi) {   // math32.go:38-45
  if (v < lo) return lo;
  else if (v > hi) return hi;
  return v;
}

// Go's int(float32) on amd64: truncation; NaN / out of range give min int64.
__device__ __forceinline__ int64_t go_int(float f) {
  if (!(f >= -9.2233720368547758e18f && f < 9.2233720368547758e18f)) return INT64_MIN;
  return (int64_t)f;
}

__device__ __forceinline__ V3 ld3(const float* __restrict__ p, uint32_t i) {
  return v3(p[3 * (size_t)i + 0], p[3 * (size_t)i + 1], p[3 * (size_t)i + 2]);
}

// ---------------------------------------------------------------------------
// lighting
// ---------------------------------------------------------------------------

// LightVolume cell lookup: convertToXYZIndices + getLightCell (light.go:717-724, 1267-1272);
// the cell's index, or -1 where there is none.
__device__ __forceinline__ int64_t volume_cell_index(const LightGPU& l, V3 p) {
  int64_t x = go_int(roundf(fdiv(p.x - l.offset[0] - l.vmin[0] - fdiv(l.cell[0], 2.0f), l.cell[0])));
  int64_t y = go_int(roundf(fdiv(p.y - l.offset[1] - l.vmin[1] - fdiv(l.cell[1], 2.0f), l.cell[1])));
  int64_t z = go_int(roundf(fdiv(p.z - l.offset[2] - l.vmin[2] - fdiv(l.cell[2], 2.0f), l.cell[2])));
  if (x < 0 || y < 0 || z < 0 || x >= l.cx || y >= l.cy || z >= l.cz) return -1;
  int64_t i = x + (y * l.cx) + (z * l.cx * l.cy);
  if (!l.cells || i >= (int64_t)l.ncells) return -1;
  return i;
}

__device__ __forceinline__ const float* volume_cell(const LightGPU& l, V3 p) {
  const int64_t i = volume_cell_index(l, p);
  return i < 0 ? nullptr : l.cells + 4 * i;
}

// VertexLight of an ambient, point or directional light on one vertex
// (light.go:116-120, 191-258, 332-360), accumulating into rgb in the
// reference's order.
__device__ __forceinline__ void light_direct(const LightGPU& l, const PartLight& pl, const PartGPU& P, uint32_t flags,
                                             V3 vp, V3 vn, float& r, float& g, float& b) {
  const uint32_t lmode = (flags >> PF_LMODE_SHIFT) & 3u;
  if (l.type == T3D_LIGHT_AMBIENT) {
    r += l.amb[0];
    g += l.amb[1];
    b += l.amb[2];
  } else if (l.type == T3D_LIGHT_POINT) {
    V3 wp = v3(pl.v[0], pl.v[1], pl.v[2]);
    float distance = dist_sq(wp, vp);
    if (l.range > 0) {
      if (distance > l.range_sq) return;
    } else {
      if (fdiv(1.0f, distance) * l.energy < 0.001f) return;
    }
    V3 lv = unit(sub(wp, vp));
    V3 n = lmode == T3D_LIGHTING_FIXED_NORMALS ? lv : vn;
    float diffuse = dot(n, lv);
    if (lmode == T3D_LIGHTING_DOUBLE_SIDED) diffuse = fabsf(diffuse);
    if (diffuse > 0) {
      float df = diffuse * fdiv(1.0f, 1.0f + (0.1f * distance)) * 2.0f;
      if (l.range_sq > 0) {
        float dc = clampf(fdiv(l.range_sq - distance, distance), 0.0f, 1.0f);
        df *= dc;
      }
      r += l.color[0] * df * l.energy;
      g += l.color[1] * df * l.energy;
      b += l.color[2] * df * l.energy;
    }
  } else if (l.type == T3D_LIGHT_DIRECTIONAL) {
    V3 fwd = v3(l.fwd[0], l.fwd[1], l.fwd[2]);
    V3 n;
    if (lmode == T3D_LIGHTING_FIXED_NORMALS) n = fwd;
    else if (flags & PF_SKINNED) n = vn;
    else n = mult_vec(P.rd, vn);
    float df = dot(n, fwd);
    if (lmode == T3D_LIGHTING_DOUBLE_SIDED) df = fabsf(df);
    if (df <= 0) return;
    r += l.color[0] * df * l.energy;
    g += l.color[1] * df * l.energy;
    b += l.color[2] * df * l.energy;
  }
}

// One light of a pair's list on one vertex: light_direct, or a LightVolume
// (light.go:583-663), whose cell first runs the lights its useLights names
// (light.go:650-653; entries ls/pls[nest_off + j] of the same list) and then
// adds its own colour.
__device__ __forceinline__ void light_vertex(const LightGPU& l, const PartLight& pl, const LightGPU* ls,
                                             const PartLight* pls, const PartGPU& P, uint32_t flags, V3 vp, V3 vn,
                                             float& r, float& g, float& b) {
  if (l.type != T3D_LIGHT_VOLUME) {
    light_direct(l, pl, P, flags, vp, vn, r, g, b);
  } else {
    const uint32_t mode = (flags >> PF_LVMODE_SHIFT) & 3u;
    if (mode == T3D_LV_OFF) return;                                // LightVolume.Light, light.go:671-673
    float c0 = 0, c1 = 0, c2 = 0;
    if (mode == T3D_LV_PER_OBJECT) {
      if (pl.cell_found) {
        c0 = pl.v[0];
        c1 = pl.v[1];
        c2 = pl.v[2];
      }
    } else {
      V3 halfc = v3(fdiv(l.cell[0], 2.0f), fdiv(l.cell[1], 2.0f), fdiv(l.cell[2], 2.0f));
      V3 bias = v3(l.bias[0], l.bias[1], l.bias[2]);
      V3 vpos = v3(0, 0, 0), vpn = v3(0, 0, 0);
      if (flags & PF_SKINNED) {
        if (mode == T3D_LV_PER_VERTEX_WITH_NORMAL) vpn = add(add(vpos, hmul(vn, halfc)), bias);
        else vpos = add(vp, bias);
      } else {
        if (mode == T3D_LV_PER_VERTEX_WITH_NORMAL) vpn = add(mult_vec(P.model, add(vp, hmul(vn, halfc))), bias);
        else vpos = add(mult_vec(P.model, vp), bias);
      }
      if (!(vpos.x >= l.vmin[0] && vpos.y >= l.vmin[1] && vpos.z >= l.vmin[2] && vpos.x <= l.vmax[0] &&
            vpos.y <= l.vmax[1] && vpos.z <= l.vmax[2]))
        return;                                                      // Dimensions.Contains, mesh.go:80-82
      const int64_t ci = volume_cell_index(l, mode == T3D_LV_PER_VERTEX_WITH_NORMAL ? vpn : vpos);
      if (ci >= 0) {
        const float* cell = l.cells + 4 * ci;
        c0 = cell[0];
        c1 = cell[1];
        c2 = cell[2];
        if (l.cell_lights) {
          const uint32_t j1 = l.cell_lights[ci + 1];
          const uint32_t* ix = l.cell_lights + l.ncells + 1;
          for (uint32_t j = l.cell_lights[ci]; j < j1; j++) {
            const uint32_t li = l.nest_off + ix[j];
            light_direct(ls[li], pls[li], P, flags, vp, vn, r, g, b);
          }
        }
      }
    }
    r += l.color[0] * c0 * l.energy;
    g += l.color[1] * c1 * l.energy;
    b += l.color[2] * c2 * l.energy;
  }
}

// ---------------------------------------------------------------------------
// k_vertex
// ---------------------------------------------------------------------------

__global__ __launch_bounds__(256) void k_vertex(FrameArgs a, uint32_t blk0) {
  const uint32_t lb = frame_block(a, FC_VCHUNKS, 1);
  if (lb == kNoBlock) return;
  const uint32_t b = blk0 + lb;
  const uint32_t p = a.vchunk_part[b];
  const PartGPU P = a.parts[p];        // by value: read once, before any store
  const uint32_t flags = P.flags;

  __shared__ LightGPU s_light[kMaxLights];
  __shared__ PartLight s_pl[kMaxLights];
  // bone palette rows in LDS, five float4 per bone (four rows and one of
  // padding): lanes blending different bones then start on different LDS
  // banks instead of two bank groups (a 64-byte stride covers 16 of 32 banks)
  constexpr uint32_t kBoneRows = 5;
  __shared__ float4 s_bone_rows[kMaxLdsBones * kBoneRows];

  const uint32_t nl = (flags & PF_LIGHTING) ? P.light_count : 0u;
  const bool lds_lights = nl <= kMaxLights;   // more lights are read from device memory as the loop needs them
  if (lds_lights)
    for (uint32_t i = threadIdx.x; i < nl; i += blockDim.x) {
      s_pl[i] = a.plights[P.light_base + i];
      s_light[i] = a.lights[s_pl[i].light];
    }
  const bool skinned = (flags & PF_SKINNED) != 0;
  const bool lds_bones = skinned && P.bone_count <= kMaxLdsBones;
  if (lds_bones) {
    const float4* src = reinterpret_cast<const float4*>(P.palette);
    for (uint32_t i = threadIdx.x; i < 4u * P.bone_count; i += blockDim.x) s_bone_rows[(i / 4) * kBoneRows + i % 4] = src[i];
  }
  __shared__ M4 s_umvp;                  // unaltered MVP of DepthModeUnbillboarded (model.go:697-703)
  if ((flags & PF_UNBB) && threadIdx.x == 0) s_umvp = mult(P.model, a.fc[P.cam].vp);
  __syncthreads();

  const FrameConst fc = a.fc[P.cam];
  const uint32_t first = (b - a.vchunk_prefix[p]) * a.vchunk;
  const uint32_t last = min(first + a.vchunk, P.vcount);
  for (uint32_t i = first + threadIdx.x; i < last; i += blockDim.x) {
    const uint32_t v = P.vmin + i;
    // VertexPositionWithShapeKeys / VertexNormalWithShapeKeys (mesh.go:796-822)
    V3 pos = ld3(P.pos, v);
    V3 nrm = ld3(P.nrm, v);
    if (flags & PF_SHAPE_KEYS) {
      for (uint32_t k = 0; k < P.shape_keys; k++) {
        const float w = P.sk_w[k];
        const size_t o = ((size_t)k * P.mesh_vcount + v) * 3;
        pos.x += P.sk_pos[o + 0] * w;
        pos.y += P.sk_pos[o + 1] * w;
        pos.z += P.sk_pos[o + 2] * w;
      }
      for (uint32_t k = 0; k < P.shape_keys; k++) {
        const float w = P.sk_w[k];
        const size_t o = ((size_t)k * P.mesh_vcount + v) * 3;
        nrm.x += P.sk_nrm[o + 0] * w;
        nrm.y += P.sk_nrm[o + 1] * w;
        nrm.z += P.sk_nrm[o + 2] * w;
      }
    }
    V3 vp, vn;
    V4 clip;
    if (skinned) {
      // Model.skinVertex (model.go:559-596)
      const ushort4 bi = reinterpret_cast<const ushort4*>(P.bones)[v];
      const float4 bw = reinterpret_cast<const float4*>(P.weights)[v];
      const uint16_t bidx[4] = {bi.x, bi.y, bi.z, bi.w};
      const float bwt[4] = {bw.x, bw.y, bw.z, bw.w};
      M4 S = zero4();
      for (int k = 0; k < 4; k++) {
        const float w = bwt[k];
        if (w == 0) continue;
        float4 B[4];
        for (int r = 0; r < 4; r++)
          B[r] = lds_bones ? s_bone_rows[bidx[k] * kBoneRows + r]
                           : reinterpret_cast<const float4*>(P.palette)[(size_t)bidx[k] * 4 + r];
        if (w == 1) {
          for (int r = 0; r < 4; r++) {
            S.m[r][0] = B[r].x;
            S.m[r][1] = B[r].y;
            S.m[r][2] = B[r].z;
            S.m[r][3] = B[r].w;
          }
          break;
        }
        for (int r = 0; r < 4; r++) {
          S.m[r][0] += B[r].x * w;
          S.m[r][1] += B[r].y * w;
          S.m[r][2] += B[r].z * w;
          S.m[r][3] += B[r].w * w;
        }
      }
      vp = mult_vec(S, pos);
      S.m[3][0] = 0;
      S.m[3][1] = 0;
      S.m[3][2] = 0;
      S.m[3][3] = 1;
      vn = mult_vec(S, nrm);
      clip = mult_vec_w(fc.vp, vp);                                 // model.go:772-775
    } else {
      vp = pos;
      vn = nrm;
      clip = mult_vec_w(P.mvp, pos);                                // model.go:795-798
      if (flags & PF_UNBB) a.vud[P.vbase + i] = mult_vec_w(s_umvp, pos).z;   // model.go:800-803, not snapped
    }
    if (fc.snap > 0) {                                              // model.go:807-812
      clip.x = fdiv(roundf(clip.x * fc.snap), fc.snap);
      clip.y = fdiv(roundf(clip.y * fc.snap), fc.snap);
      clip.z = fdiv(roundf(clip.z * fc.snap), fc.snap);
    }
    // vertex colour (camera.go:1750-1768)
    float cr, cg, cb, ca;
    if (flags & PF_HAS_COLOR) {
      const float4 vc = reinterpret_cast<const float4*>(P.col)[v];
      cr = vc.x * P.mp[0];
      cg = vc.y * P.mp[1];
      cb = vc.z * P.mp[2];
      ca = vc.w * P.mp[3];
    } else {
      cr = P.mp[0];
      cg = P.mp[1];
      cb = P.mp[2];
      ca = P.mp[3];
    }
    if (flags & PF_LIGHTING) {
      // vertexLights reset to (0,0,0,1) then every light in order (camera.go:1613-1641)
      float lr = 0, lg = 0, lb = 0;
      if (lds_lights) {
        for (uint32_t k = 0; k < nl; k++) {
          if (!s_pl[k].enabled) continue;
          light_vertex(s_light[k], s_pl[k], s_light, s_pl, P, flags, vp, vn, lr, lg, lb);
        }
      } else {
        const LightGPU* ls = a.lights + P.light0;
        const PartLight* pls = a.plights + P.light_base;
        for (uint32_t k = 0; k < nl; k++) {
          const PartLight pl = pls[k];
          if (!pl.enabled) continue;
          light_vertex(ls[k], pl, ls, pls, P, flags, vp, vn, lr, lg, lb);
        }
      }
      cr *= lr;
      cg *= lg;
      cb *= lb;
    }
    if (fc.fog_active) {                                            // camera.go:1786-1810
      float depth = fdiv(clip.z + fc.dmp, fc.spread);
      if (depth < 0) depth = 0;
      else if (depth > 1) depth = 1;
      depth = fc.fog_range[0] + ((fc.fog_range[1] - fc.fog_range[0]) * 1.0f - depth);
      if (fc.fog_mode == T3D_FOG_ADD) {
        cr += fc.fog_color[0] * depth;
        cg += fc.fog_color[1] * depth;
        cb += fc.fog_color[2] * depth;
      } else if (fc.fog_mode == T3D_FOG_SUB) {
        cr *= fc.fog_color[0] * depth;
        cg *= fc.fog_color[1] * depth;
        cb *= fc.fog_color[2] * depth;
      }
    }
    const uint32_t o = P.vbase + i;
    if (fc.render_normals) {   // model.go:814-818: the rotation applied to the position; colour camera.go:1718-1720
      const V3 tn = mult_vec(P.nrot, vp);
      a.vnrm[o] = make_float4(tn.x * 0.5f + 0.5f, tn.y * 0.5f + 0.5f, tn.z * 0.5f + 0.5f, 0.0f);
    }
    a.vclip[(size_t)o * a.vstride] = make_float4(clip.x, clip.y, clip.z, clip.w);
    a.vcol[(size_t)o * a.vstride] = make_float4(cr, cg, cb, ca);
    if (skinned) a.vpos[o] = make_float4(vp.x, vp.y, vp.z, 0.0f);
  }
}

// ---------------------------------------------------------------------------
// k_palettes, k_uniforms
// ---------------------------------------------------------------------------

// k_palettes: the bone palettes of armatures given as poses, one block per
// armature, computed as Node.Transform() does down the armature
// (node.go:572-608):
//   1. Scale·rotation·Translate of every bone, spread over the block;
//   2. Transform() = local·parent's Transform() (the armature parent's for a
//      root bone), level by level: each round, every bone whose parent is
//      done takes its parent's product, so the rounds number the armature's
//      depth, not its bones;
//   3. boneInfluence = inverseBindMatrix·Transform() (node.go:592-594).
// Every product is Matrix4.Mult in its own order, so the palette equals the
// host's bit for bit.  Parents and round states live in LDS; the bone
// transforms too up to kLdsPoseBones, in device memory past that.
constexpr uint32_t kLdsPoseBones = 256;     // 16 KB of LDS
constexpr uint32_t kLevelPoseBones = 4096;  // armatures walked level by level up to this size
constexpr uint32_t kPoseThreads = 64;

// Local transform of bone b: Scale·rotation·Translate of its pose, after the
// armature's animation (if any) has set the channels it samples.
__device__ __forceinline__ M4 bone_local(const PaletteJob& J, uint32_t b) {
  const t3d_bone_pose& B = J.poses[b];
  float pos[3] = {B.position[0], B.position[1], B.position[2]};
  float scl[3] = {B.scale[0], B.scale[1], B.scale[2]};
  M4 rot = B.rotation;
  if (b < J.nchannels) animate_bone(J.channels[b], J.time, pos, scl, rot);
  return mult(mult(scale4(scl[0], scl[1], scl[2]), rot), translate(pos[0], pos[1], pos[2]));
}

template <typename WPtr>
__device__ __forceinline__ void pose_armature(const PaletteJob& J, WPtr W, int16_t* par, uint8_t* state) {
  const uint32_t n = J.n;
  for (uint32_t b = threadIdx.x; b < n; b += kPoseThreads) {
    W[b] = bone_local(J, b);
    par[b] = (int16_t)J.poses[b].parent;            // -1 <= parent < b, checked by the host
    state[b] = 0;                                    // 0 waiting, 1 ready this round, 2 done
  }
  __syncthreads();
  // Each round readies at least the lowest waiting bone (its parent is earlier, so done): at most n rounds.
  // A parent index outside [-1, b) (refused by the host) never readies, and the bound still ends the loop.
  for (uint32_t round = 0; round < n; round++) {
    bool waiting = false;
    for (uint32_t b = threadIdx.x; b < n; b += kPoseThreads)
      if (state[b] == 0) {
        const int32_t p = par[b];
        if (p < 0 || ((uint32_t)p < b && state[p] == 2)) state[b] = 1;   // a parent going 0 -> 1 now reads as not done
        else waiting = true;
      }
    __syncthreads();
    for (uint32_t b = threadIdx.x; b < n; b += kPoseThreads)
      if (state[b] == 1) {
        const int32_t p = par[b];
        if (p >= 0) W[b] = mult(W[b], W[p]);         // W[p] was finished in an earlier round
        else if (J.root) W[b] = mult(W[b], *J.root);
        state[b] = 2;
      }
    if (!__syncthreads_or(waiting)) break;
  }
  for (uint32_t b = threadIdx.x; b < n; b += kPoseThreads) {
    const M4 r = mult(J.poses[b].inverse_bind, W[b]);
    float4* out = reinterpret_cast<float4*>(J.palette + 16ull * b);
    for (int i = 0; i < 4; i++) out[i] = make_float4(r.m[i][0], r.m[i][1], r.m[i][2], r.m[i][3]);
  }
}

// Armatures past kLevelPoseBones: the parent chain walked in bone order by
// one thread (a parent is always an earlier bone).
__device__ void pose_armature_serial(const PaletteJob& J, M4* W) {
  for (uint32_t b = threadIdx.x; b < J.n; b += kPoseThreads) W[b] = bone_local(J, b);
  __syncthreads();
  if (threadIdx.x == 0)
    for (uint32_t b = 0; b < J.n; b++) {
      const int32_t p = J.poses[b].parent;
      if (p >= 0 && (uint32_t)p < b) W[b] = mult(W[b], W[p]);
      else if (J.root) W[b] = mult(W[b], *J.root);
    }
  __syncthreads();
  for (uint32_t b = threadIdx.x; b < J.n; b += kPoseThreads) {
    const M4 r = mult(J.poses[b].inverse_bind, W[b]);
    float4* out = reinterpret_cast<float4*>(J.palette + 16ull * b);
    for (int i = 0; i < 4; i++) out[i] = make_float4(r.m[i][0], r.m[i][1], r.m[i][2], r.m[i][3]);
  }
}

__global__ __launch_bounds__(kPoseThreads) void k_palettes(FrameArgs a) {
  const PaletteJob J = a.pal_jobs[blockIdx.x];
  __shared__ M4 s_world[kLdsPoseBones];
  __shared__ int16_t s_par[kLevelPoseBones];
  __shared__ uint8_t s_state[kLevelPoseBones];
  M4* gw = a.pose_world + J.world_off;
  if (J.n <= kLdsPoseBones) pose_armature(J, s_world, s_par, s_state);
  else if (J.n <= kLevelPoseBones) pose_armature(J, gw, s_par, s_state);
  else pose_armature_serial(J, gw);
}

// Per-pair uniforms: MVP (model.go:695), invertedCamPos (model.go:629-630),
// the directional working rotation (light.go:326-330) and, per light of the
// pair, beginModel's working position (light.go:175-189), the per-model
// point-light skip (camera.go:1592, 1625-1629) and the per-object LightVolume
// cell (light.go:610-628).  The same float32 functions the host would run, so
// the values are the ones the reference computes on the CPU.  One wavefront
// per render pair: lane k handles the pair's lights k, k + 64, ..., lane 0 also the pair's own uniforms and the reset of its depth range
// and survivor count, so the latency is one light's, not the sum over lights;
// all lanes first write the pair's entries of the block->pair maps.
__global__ __launch_bounds__(64) void k_uniforms(FrameArgs a, uint32_t part0, uint32_t nparts) {
  if (past_count(a, FC_PAIRS)) return;
  PartGPU& P = const_cast<PartGPU*>(a.parts)[part0 + blockIdx.x];
  if (a.counts) {   // a device-built frame: the pair's record is built here from the model, mesh and material
    __shared__ PartGPU s_part;
    if (threadIdx.x == 0) build_pair(a, part0 + blockIdx.x, s_part);
    __syncthreads();
    static_assert(sizeof(PartGPU) % 16 == 0, "whole float4 rows");
    const float4* src = reinterpret_cast<const float4*>(&s_part);
    float4* dst = reinterpret_cast<float4*>(&P);
    for (uint32_t i = threadIdx.x; i < sizeof(PartGPU) / 16; i += 64) dst[i] = src[i];
    __syncthreads();
  }
  const uint32_t k = threadIdx.x;
  {  // block -> pair maps of the later kernels: one load where a search over the prefixes took nine
    const uint32_t p = part0 + blockIdx.x;
    for (uint32_t i = a.tile_prefix[p] + k; i < a.tile_prefix[p + 1]; i += 64) a.tile_part[i] = p;
    for (uint32_t i = a.vchunk_prefix[p] + k; i < a.vchunk_prefix[p + 1]; i += 64) a.vchunk_part[i] = p;
  }
  const uint32_t flags = P.flags;
  const uint32_t nl = (flags & PF_LIGHTING) ? P.light_count : 0u;
  if (k > 0 && k >= nl) return;
  const FrameConst& fc = a.fc[P.cam];
  const M4 M = P.model;
  V3 dp, ds;
  M4 dr;
  decompose(inverted(M), dp, ds, dr);
  const V3 is = v3(fdiv(1.0f, ds.x), fdiv(1.0f, ds.y), fdiv(1.0f, ds.z));
  for (uint32_t kk = k; kk < nl; kk += 64) {
    const bool skinned = (flags & PF_SKINNED) != 0;
    const uint32_t lvmode = (flags >> PF_LVMODE_SHIFT) & 3u;
    const V3 modelPos = row3(M);
    PartLight pl{};
    pl.light = P.light0 + kk;
    const LightGPU& L = a.lights[pl.light];
    pl.enabled = L.nested ? 0u : 1u;   // a volume's cell light runs from its cells only, unskipped
    if (L.type == T3D_LIGHT_POINT) {
      const V3 wp = v3(L.pos[0], L.pos[1], L.pos[2]);
      const V3 w = skinned ? wp : add(mult_vec(dr, wp), hmul(dp, is));
      pl.v[0] = w.x;
      pl.v[1] = w.y;
      pl.v[2] = w.z;
      if (L.range > 0 && !L.nested) {
        const float dist = P.max_span + L.range;
        if (dist_sq(modelPos, wp) > dist * dist) pl.enabled = 0;
      }
    } else if (L.type == T3D_LIGHT_VOLUME && lvmode == T3D_LV_PER_OBJECT) {
      const float* cell = volume_cell(L, modelPos);
      if (cell) {
        pl.cell_found = 1;
        pl.v[0] = cell[0];
        pl.v[1] = cell[1];
        pl.v[2] = cell[2];
      }
    }
    a.plights[P.light_base + kk] = pl;
  }
  if (k == 0) {
    uint32_t* st = a.stats + 4 * (part0 + blockIdx.x);
    st[0] = __float_as_uint(kMaxF32);  // minDepth = MaxFloat32 (model.go:702)
    st[1] = 0u;                        // max over non-negative keys
    st[2] = 0u;                        // survivors
    st[3] = 0u;
    P.mvp = mult(P.mvp, fc.vp);          // the uploaded base (model.go:634-695)
    if (fc.render_normals) {             // mvJustRForNormals (model.go:705-708)
      V3 np, ns;
      M4 nr;
      decompose(mult(M, fc.view), np, ns, nr);
      P.nrot = nr;
    }
    const V3 icp = add(mult_vec(dr, v3(fc.campos[0], fc.campos[1], fc.campos[2])), hmul(dp, is));
    P.icp[0] = icp.x;
    P.icp[1] = icp.y;
    P.icp[2] = icp.z;
    if (flags & PF_NEED_RD) P.rd = transposed(inverted(P.rd));
  }
}

__device__ __forceinline__ float wave_min(float v) {
  for (int o = 32; o > 0; o >>= 1) v = fminf(v, __shfl_xor(v, o));
  return v;
}
__device__ __forceinline__ float wave_max(float v) {
  for (int o = 32; o > 0; o >>= 1) v = fmaxf(v, __shfl_xor(v, o));
  return v;
}
__device__ __forceinline__ uint32_t wave_sum(uint32_t v) {
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o);
  return v;
}

// ---------------------------------------------------------------------------
// k_cull
// ---------------------------------------------------------------------------

// kThreads: 256, 512 or 1024 threads per tile block (tunable).
template <uint32_t kThreads>
__global__ __launch_bounds__(kThreads) void k_cull(FrameArgs a, uint32_t blk0) {
  constexpr uint32_t kCullWaves = kThreads / 64;
  const uint32_t lb = frame_block(a, FC_TILES, 1);
  if (lb == kNoBlock) return;
  const uint32_t b = blk0 + lb;
  const uint32_t p = a.tile_part[b];
  const PartGPU P = a.parts[p];        // by value: read once, before any store
  const uint32_t flags = P.flags;
  const FrameConst fc = a.fc[P.cam];
  const uint32_t t0 = (b - a.tile_prefix[p]) * a.tile;
  const uint32_t t1 = min(t0 + a.tile, P.tcount);
  const bool skinned = (flags & PF_SKINNED) != 0;
  const bool backface = (flags & PF_BACKFACE) != 0;
  const bool all = (flags & PF_ALL) != 0;     // processOnlyVisible false: no test removes a triangle
  const V3 camPos = v3(fc.campos[0], fc.campos[1], fc.campos[2]);
  const V3 icp = v3(P.icp[0], P.icp[1], P.icp[2]);

  // Survivors are compacted in triangle order to the front of the tile's
  // range of key[] and tri_rec[]; tile_count[b] says how many there are.
  float mn = kMaxF32, mx = 0.0f;
  const uint32_t wv = threadIdx.x / warpSize, ln = threadIdx.x % warpSize;
  const uint64_t lt_mask = (ln == 0) ? 0ull : (~0ull >> (64 - ln));
  __shared__ uint32_t s_wc[kCullWaves];
  uint32_t run = 0;                                 // survivors so far, block-uniform
  for (uint32_t base = t0; base < t1; base += blockDim.x) {
    const uint32_t t = base + threadIdx.x;
    bool keep = false;
    float depth = -1.0f;
    uint32_t ov[3] = {0, 0, 0};
    if (t < t1) {
      uint32_t i0, i1, i2;
      if (P.idx16) {                              // uniform: the pair's mesh index width
        i0 = P.idx16[3 * (size_t)t + 0];
        i1 = P.idx16[3 * (size_t)t + 1];
        i2 = P.idx16[3 * (size_t)t + 2];
      } else {
        i0 = P.idx[3 * (size_t)t + 0];
        i1 = P.idx[3 * (size_t)t + 1];
        i2 = P.idx[3 * (size_t)t + 2];
      }
      const uint32_t vi[3] = {i0, i1, i2};
      float nx[3], ny[3];
      bool outOfBounds = true;
      V3 stc = v3(0, 0, 0);
      // every gather of the triangle first, so their latencies overlap: clip
      // positions, and what its depth key needs (a static triangle's stored
      // Triangle.Center, or the three skinned positions)
      float4 cc[3];
      V3 q[3];
      for (int k = 0; k < 3; k++) {
        ov[k] = P.vbase + (vi[k] - P.vmin);
        cc[k] = a.vclip[(size_t)ov[k] * a.vstride];
      }
      if (skinned) {
        for (int k = 0; k < 3; k++) {
          const float4 s = a.vpos[ov[k]];
          q[k] = v3(s.x, s.y, s.z);
        }
      } else {
        const float* cp = P.ctr + 3 * (size_t)t;
        q[0] = v3(cp[0], cp[1], cp[2]);
      }
      for (int k = 0; k < 3; k++) {                                   // model.go:855-886
        const float4 c = cc[k];
        if (skinned) {
          stc.x += q[k].x;
          stc.y += q[k].y;
          stc.z += q[k].z;
        }
        float w = c.w;
        if (w < 0) w = 0.000001f;
        nx[k] = fdiv(c.x, w);
        ny[k] = fdiv(c.y, w);
        if (c.z + 1.0f >= fc.near_ && c.z < fc.far_) outOfBounds = false;
      }
      keep = all || !outOfBounds;
      if (keep && !all && ((nx[0] < -0.5f && nx[1] < -0.5f && nx[2] < -0.5f) ||      // model.go:890-898
                   (nx[0] > 0.5f && nx[1] > 0.5f && nx[2] > 0.5f) ||
                   (ny[0] < -0.5f && ny[1] < -0.5f && ny[2] < -0.5f) ||
                   (ny[0] > 0.5f && ny[1] > 0.5f && ny[2] > 0.5f)))
        keep = false;
      if (keep && !all && backface) {                                  // model.go:902-925
        const float n0x = nx[0] - nx[1];
        const float n0y = ny[0] - ny[1];
        const float n1x = nx[1] - nx[2];
        const float n1y = ny[1] - ny[2];
        const float nor = (n0x * n1y) - (n1x * n0y);
        if (nor < 0) keep = false;
      }
      if (keep) {                                                       // model.go:941-960
        if (skinned) {
          const float dx = camPos.x - fdiv(stc.x, 3.0f);
          const float dy = camPos.y - fdiv(stc.y, 3.0f);
          const float dz = camPos.z - fdiv(stc.z, 3.0f);
          depth = dx * dx + dy * dy + dz * dz;
        } else {
          const float dx = icp.x - q[0].x;    // Triangle.Center, stored at upload (mesh.go:1980-1982)
          const float dy = icp.y - q[0].y;
          const float dz = icp.z - q[0].z;
          depth = dx * dx + dy * dy + dz * dz;
        }
        if (!all && (isnan(depth) || isinf(depth))) {
          keep = false;
          depth = -1.0f;
        }
      }
    }
    const uint64_t m = __ballot(keep);
    if (ln == 0) s_wc[wv] = (uint32_t)__popcll(m);
    __syncthreads();
    uint32_t wbase = 0, tot = 0;
    for (uint32_t w = 0; w < kCullWaves; w++) {
      if (w < wv) wbase += s_wc[w];
      tot += s_wc[w];
    }
    if (keep) {
      const uint32_t pos = P.tbase + t0 + run + wbase + (uint32_t)__popcll(m & lt_mask);
      a.key[pos] = depth;
      a.tri_rec[pos] = make_uint4(ov[0], ov[1], ov[2], t);
      mn = fminf(mn, depth);
      mx = fmaxf(mx, depth);
    }
    run += tot;
    __syncthreads();                              // s_wc is rewritten next round
  }
  // block reduction, one atomic of each kind per block
  __shared__ float s_mn[kCullWaves], s_mx[kCullWaves];
  mn = wave_min(mn);
  mx = wave_max(mx);
  if (ln == 0) {
    s_mn[wv] = mn;
    s_mx[wv] = mx;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    for (uint32_t w = 1; w < kCullWaves; w++) {
      mn = fminf(mn, s_mn[w]);
      mx = fmaxf(mx, s_mx[w]);
    }
    a.tile_count[b] = run;
    if (run) {
      atomicMin(&a.stats[4 * p + 0], __float_as_uint(mn));   // keys are >= +0: uint order = float order
      atomicMax(&a.stats[4 * p + 1], __float_as_uint(mx));
      atomicAdd(&a.stats[4 * p + 2], run);
    }
  }
}

// ---------------------------------------------------------------------------
// sort: bins, histograms, scan, stable scatter
// ---------------------------------------------------------------------------

// Bin of a survivor in visit order: sortingTriangleBucket.Sort's
// int(clamp(((d - min) / (max - min)) * 512, 0, 511)) (sortingTriangles.go:38-64),
// reversed for BackToFront so ascending keys follow ForEach (sortingTriangles.go:93-125).
__device__ __forceinline__ uint32_t visit_bin(float d, float mn, float rd, uint32_t sort_mode) {
  int64_t bin = 0;
  if (sort_mode != T3D_SORT_NONE) {
    const float x = fdiv(d - mn, rd) * (float)kBins;
    bin = go_int(clampf(x, 0.0f, (float)(kBins - 1)));
  }
  return sort_mode == T3D_SORT_BACK_TO_FRONT ? (kBins - 1) - (uint32_t)bin : (uint32_t)bin;
}

// Lanes of `valid` (a ballot) whose 9-bit bin equals this lane's: one
// ballot per bit, so the cost does not depend on how many distinct bins the
// wavefront holds.  Meaningful for valid lanes only.
__device__ __forceinline__ uint64_t bin_peers(uint32_t bin, uint64_t valid) {
  static_assert(kBins == 512, "nine bin bits");
  uint64_t peers = valid;
  for (int bit = 0; bit < 9; bit++) {
    const bool set = (bin >> bit) & 1u;
    const uint64_t m = __ballot(set);
    peers &= set ? m : ~m;
  }
  return peers;
}

__device__ __forceinline__ void part_range(const uint32_t* stats, uint32_t p, float& mn, float& rd) {
  mn = __uint_as_float(stats[4 * p + 0]);
  const float mx = __uint_as_float(stats[4 * p + 1]);
  rd = mx - mn;
  if (rd == 0) rd += 0.001f;                                      // sortingTriangles.go:43-46
}

constexpr uint32_t kHistThreads = 256;   // four wavefronts: a 2,048-triangle tile's keys in one round of loads
__global__ __launch_bounds__(kHistThreads) void k_hist(FrameArgs a, uint32_t blk0) {
  const uint32_t lb = frame_block(a, FC_TILES, 1);
  if (lb == kNoBlock) return;
  const uint32_t b = blk0 + lb;
  const uint32_t p = a.tile_part[b];
  const PartGPU P = a.parts[p];        // by value: read once, before any store
  if (P.ntiles <= 1) return;
  __shared__ uint32_t h[kBins];
  for (uint32_t i = threadIdx.x; i < kBins; i += blockDim.x) h[i] = 0;
  __syncthreads();
  const uint32_t tl = b - a.tile_prefix[p];
  const uint32_t n = a.tile_count[b];                // survivors, compacted at the tile's front
  const float* __restrict__ keys = a.key + P.tbase + tl * a.tile;
  float mn, rd;
  part_range(a.stats, p, mn, rd);
  const uint32_t mode = (P.flags >> PF_SORT_SHIFT) & 3u;
  for (uint32_t ib = 0; ib < n; ib += 8 * kHistThreads) {     // eight independent loads in flight per lane
    float d[8];
    for (int j = 0; j < 8; j++) {
      const uint32_t i = ib + j * kHistThreads + threadIdx.x;
      d[j] = i < n ? keys[i] : 0.0f;
    }
    // every survivor counts, a NaN key (kept when processOnlyVisible is false) included
    for (int j = 0; j < 8; j++)
      if (ib + j * kHistThreads + threadIdx.x < n) atomicAdd(&h[visit_bin(d[j], mn, rd, mode)], 1u);
  }
  __syncthreads();
  uint32_t* col = a.hist + P.hbase + tl;
  for (uint32_t i = threadIdx.x; i < kBins; i += blockDim.x) col[(size_t)i * P.ntiles] = h[i];
}

// Scan of each multi-tile part's histogram matrix, laid out bin-major
// (element bin*ntiles + tile), which is exactly the output order.
__device__ __forceinline__ uint32_t block_exclusive_scan_256(uint32_t v, uint32_t* s_tmp, uint32_t& total) {
  const uint32_t ln = threadIdx.x % warpSize, wv = threadIdx.x / warpSize;
  uint32_t x = v;
  for (int o = 1; o < 64; o <<= 1) {
    const uint32_t y = __shfl_up(x, o);
    if (ln >= (uint32_t)o) x += y;
  }
  if (ln == 63) s_tmp[wv] = x;
  __syncthreads();
  uint32_t wbase = 0, tot = 0;
  for (uint32_t w = 0; w < blockDim.x / 64; w++) {
    if (w < wv) wbase += s_tmp[w];
    tot += s_tmp[w];
  }
  __syncthreads();
  total = tot;
  return wbase + x - v;
}

__global__ __launch_bounds__(256) void k_scan_reduce(FrameArgs a, uint32_t sblk0) {
  if (past_count(a, FC_SBLOCKS)) return;
  const uint32_t b = sblk0 + blockIdx.x;
  const uint32_t p = find_part(a.sblock_prefix, a.nparts, b);
  const PartGPU P = a.parts[p];        // by value: read once, before any store
  const uint32_t len = kBins * P.ntiles;
  const uint32_t e0 = (b - a.sblock_prefix[p]) * kScanBlock;
  uint32_t s = 0;
  for (uint32_t e = e0 + threadIdx.x; e < min(e0 + kScanBlock, len); e += blockDim.x) s += a.hist[P.hbase + e];
  __shared__ uint32_t s_tmp[8];
  s = wave_sum(s);
  if (threadIdx.x % warpSize == 0) s_tmp[threadIdx.x / warpSize] = s;
  __syncthreads();
  if (threadIdx.x == 0) {
    uint32_t t = 0;
    for (uint32_t w = 0; w < blockDim.x / warpSize; w++) t += s_tmp[w];
    a.bsum[b] = t;
  }
}

// Exclusive scan of each part's scan-block sums, one 256-thread block per
// part, 256 sums per step with the running total carried across steps (a
// serial walk would pay one dependent memory round trip per scan block).
__global__ __launch_bounds__(256) void k_scan_bsums(FrameArgs a, uint32_t part0) {
  if (past_count(a, FC_PAIRS)) return;
  const uint32_t n = a.parts[part0 + blockIdx.x].nsblocks;
  if (n == 0) return;
  uint32_t* __restrict__ bs = a.bsum + a.parts[part0 + blockIdx.x].sblock0;
  __shared__ uint32_t s_tmp[8];
  uint32_t carry = 0;
  for (uint32_t c = 0; c < n; c += 256) {
    const uint32_t k = c + threadIdx.x;
    const uint32_t v = k < n ? bs[k] : 0u;
    uint32_t total;
    const uint32_t ex = block_exclusive_scan_256(v, s_tmp, total);
    if (k < n) bs[k] = carry + ex;
    carry += total;
  }
}

__global__ __launch_bounds__(256) void k_scan_down(FrameArgs a, uint32_t sblk0) {
  if (past_count(a, FC_SBLOCKS)) return;
  const uint32_t b = sblk0 + blockIdx.x;
  const uint32_t p = find_part(a.sblock_prefix, a.nparts, b);
  const PartGPU P = a.parts[p];        // by value: read once, before any store
  const uint32_t len = kBins * P.ntiles;
  const uint32_t e0 = (b - a.sblock_prefix[p]) * kScanBlock;
  __shared__ uint32_t s_tmp[8];
  uint32_t base = a.bsum[b];
  constexpr uint32_t per = kScanBlock / 256;
  uint32_t v[per];
  uint32_t local = 0;
  const uint32_t my0 = e0 + threadIdx.x * per;
  for (uint32_t k = 0; k < per; k++) {
    const uint32_t e = my0 + k;
    v[k] = e < len ? a.hist[P.hbase + e] : 0u;
    local += v[k];
  }
  uint32_t total;
  uint32_t ex = block_exclusive_scan_256(local, s_tmp, total) + base;
  for (uint32_t k = 0; k < per; k++) {
    const uint32_t e = my0 + k;
    if (e < len) a.hist[P.hbase + e] = ex;
    ex += v[k];
  }
}

// Exclusive scan of a whole [bin][tile] matrix by one 1024-thread block: each
// thread loads its contiguous segment (at most kPer 16-byte words, all loads
// issued before the first use), one block-wide scan of the 1024 segment sums,
// then each thread writes its segment back from registers.  One launch, two
// barriers and one memory round trip each way replace the three launches of
// the general scan for the (common) parts whose matrix has at most
// kScanPartMax entries.
__global__ __launch_bounds__(1024) void k_scan_part(FrameArgs a, uint32_t blk0) {
  constexpr uint32_t kPer = kScanPartMax / 4 / 1024;
  static_assert(kScanPartMax % 4096 == 0, "whole uint4 segments");
  if (past_count(a, FC_PSCAN)) return;
  const uint32_t b = blk0 + blockIdx.x;
  // the pair: a binary search of the prefixes, staged in LDS first when they fit (one memory round trip
  // instead of one per step)
  constexpr uint32_t kPrefixLds = 4096;
  __shared__ uint32_t s_pre[kPrefixLds];
  uint32_t p;
  if (a.nparts < kPrefixLds) {
    for (uint32_t i = threadIdx.x; i <= a.nparts; i += blockDim.x) s_pre[i] = a.pscan_prefix[i];
    __syncthreads();
    p = find_part(s_pre, a.nparts, b);
  } else {
    p = find_part(a.pscan_prefix, a.nparts, b);
  }
  const uint32_t n4 = kBins * a.parts[p].ntiles / 4;           // the matrix holds a multiple of 512 entries
  uint4* __restrict__ h = reinterpret_cast<uint4*>(a.hist + a.parts[p].hbase);   // 2 KB aligned
  const uint32_t tid = threadIdx.x, lane = tid % 64u, wv = tid / 64u;
  const uint32_t per = (n4 + 1023) / 1024;                     // <= kPer
  const uint32_t q0 = min(tid * per, n4), cnt = min(per, n4 - q0);
  uint4 v[kPer];
#pragma unroll
  for (uint32_t j = 0; j < kPer; j++) v[j] = j < cnt ? h[q0 + j] : make_uint4(0, 0, 0, 0);
  uint32_t sum = 0;
#pragma unroll
  for (uint32_t j = 0; j < kPer; j++) sum += v[j].x + v[j].y + v[j].z + v[j].w;
  __shared__ uint32_t s_w[16];
  uint32_t x = sum;
  for (int o = 1; o < 64; o <<= 1) {
    const uint32_t y = __shfl_up(x, o);
    if (lane >= (uint32_t)o) x += y;
  }
  if (lane == 63) s_w[wv] = x;
  __syncthreads();
  uint32_t run = x - sum;
  for (uint32_t w = 0; w < wv; w++) run += s_w[w];
#pragma unroll
  for (uint32_t j = 0; j < kPer; j++) {
    uint4 o;
    o.x = run;
    o.y = o.x + v[j].x;
    o.z = o.y + v[j].y;
    o.w = o.z + v[j].z;
    run = o.w + v[j].w;
    if (j < cnt) h[q0 + j] = o;
  }
}

// Stable counting sort, scatter pass.  One wavefront per tile walks its
// triangles in index order, 64 at a time (keys staged in LDS in one burst),
// ranks equal bins with ballots and gives every survivor its output slot:
// run[bin] + rank, where run[bin] comes from the scanned [bin][tile]
// histogram, or from a local scan when the part is a single tile.  This is
// sortingTriangleBucket.Sort + ForEach (sortingTriangles.go:38-125): bins
// visited 511->0 (BackToFront) or 0->511, insertion order kept within a bin.
// The survivor's record {vertex-record offsets, triangle} moves to its slot.
__global__ __launch_bounds__(64) void k_scatter(FrameArgs a, uint32_t blk0) {
  const uint32_t lb = frame_block(a, FC_TILES, 1);
  if (lb == kNoBlock) return;
  const uint32_t b = blk0 + lb;
  const uint32_t p = a.tile_part[b];
  const PartGPU P = a.parts[p];        // by value: read once, before any store
  if (a.stats[4 * p + 2] == 0) return;
  const uint32_t tl = b - a.tile_prefix[p];
  const uint32_t n = a.tile_count[b];                // survivors, compacted at the tile's front
  const uint32_t lane = threadIdx.x;
  const uint32_t tbase = P.tbase;
  const uint32_t mode = (P.flags >> PF_SORT_SHIFT) & 3u;
  const uint4* __restrict__ rec_in = a.tri_rec + tbase + tl * a.tile;
  uint4* __restrict__ rec_out = a.tri_vo + tbase;

  __shared__ uint32_t run[kBins];
  __shared__ float s_key[kMaxTile];
  float mn, rd;
  part_range(a.stats, p, mn, rd);
  for (uint32_t i = lane; i < n; i += 64) s_key[i] = a.key[tbase + tl * a.tile + i];   // one burst of loads

  if (P.ntiles > 1) {
    const uint32_t* col = a.hist + P.hbase + tl;
    for (uint32_t i = lane; i < kBins; i += 64) run[i] = col[(size_t)i * P.ntiles];
    __syncthreads();
  } else {
    for (uint32_t i = lane; i < kBins; i += 64) run[i] = 0;
    __syncthreads();
    for (uint32_t i = lane; i < n; i += 64) atomicAdd(&run[visit_bin(s_key[i], mn, rd, mode)], 1u);
    __syncthreads();
    uint32_t c[8], s = 0;                       // exclusive scan of 512 counters, 8 per lane
    for (int k = 0; k < 8; k++) {
      c[k] = run[lane * 8 + k];
      s += c[k];
    }
    uint32_t x = s;
    for (int o = 1; o < 64; o <<= 1) {
      const uint32_t y = __shfl_up(x, o);
      if (lane >= (uint32_t)o) x += y;
    }
    uint32_t ex = x - s;
    __syncthreads();
    for (int k = 0; k < 8; k++) {
      run[lane * 8 + k] = ex;
      ex += c[k];
    }
    __syncthreads();
  }

  const uint64_t lt_mask = (lane == 0) ? 0ull : (~0ull >> (64 - lane));
  uint4 rec_next = lane < n ? rec_in[lane] : make_uint4(0, 0, 0, 0);   // records one chunk ahead
  for (uint32_t base_i = 0; base_i < n; base_i += 64) {
    const uint32_t i = base_i + lane;
    const bool valid = i < n;
    const float d = valid ? s_key[i] : -1.0f;
    const uint64_t vmask = __ballot(valid);
    const uint4 rec = rec_next;
    if (i + 64 < n) rec_next = rec_in[i + 64];
    const uint32_t bin = valid ? visit_bin(d, mn, rd, mode) : 0xFFFFFFFFu;
    const uint64_t peers = bin_peers(bin, vmask);
    const uint32_t rank = (uint32_t)__popcll(peers & lt_mask);
    uint32_t slot = 0;
    if (valid) slot = run[bin] + rank;
    __syncthreads();
    if (valid) {
      if (rank == 0) run[bin] = slot + (uint32_t)__popcll(peers);
      rec_out[slot] = rec;
    }
    __syncthreads();
  }
}

// The same scatter with four wavefronts per tile.  Wave w owns a contiguous
// quarter of the tile's 64-triangle chunks.  A first pass over the staged keys
// counts each wave's survivors per bin; the base of wave w in bin k is
// run[k] plus the counts of waves 0..w-1 in bin k, so the four quarters rank
// independently and the result is the single-wave order exactly.
constexpr uint32_t kScatterWaves = 4;
__global__ __launch_bounds__(256) void k_scatter4(FrameArgs a, uint32_t blk0) {
  const uint32_t lb = frame_block(a, FC_TILES, 1);
  if (lb == kNoBlock) return;
  const uint32_t b = blk0 + lb;
  const uint32_t p = a.tile_part[b];
  const PartGPU P = a.parts[p];        // by value: read once, before any store
  if (a.stats[4 * p + 2] == 0) return;
  const uint32_t tl = b - a.tile_prefix[p];
  const uint32_t t0 = tl * a.tile;
  const uint32_t n = a.tile_count[b];                // survivors, compacted at the tile's front
  const uint32_t tid = threadIdx.x, lane = tid % 64u, wv = tid / 64u;
  const uint32_t tbase = P.tbase;
  const uint32_t mode = (P.flags >> PF_SORT_SHIFT) & 3u;
  const uint4* __restrict__ rec_in = a.tri_rec + tbase;
  uint4* __restrict__ rec_out = a.tri_vo + tbase;

  __shared__ uint32_t s_off[kScatterWaves][kBins];
  __shared__ uint32_t s_tmp[8];
  float mn, rd;
  part_range(a.stats, p, mn, rd);
  for (uint32_t i = tid; i < kScatterWaves * kBins; i += blockDim.x) (&s_off[0][0])[i] = 0;
  // a multi-tile pair's scanned column (each thread's two bins) is loaded now, beside the keys
  static_assert(kBins == 2 * 256, "two bins per thread");
  uint32_t col_run[2] = {0, 0};
  if (P.ntiles > 1) {
    const uint32_t* col = a.hist + P.hbase + tl;
    col_run[0] = col[(size_t)tid * P.ntiles];
    col_run[1] = col[(size_t)(tid + 256) * P.ntiles];
  }
  const uint32_t nchunks = (n + 63) / 64;
  const uint32_t c0 = (wv * nchunks) / kScatterWaves, c1 = ((wv + 1) * nchunks) / kScatterWaves;
  // every key and record of this wave's chunks (at most kChunks of 64) is
  // loaded up front, so the counting and ranking passes wait on memory once
  constexpr uint32_t kChunks = kMaxTile / 64 / kScatterWaves;
  float d[kChunks];
  uint4 rec[kChunks];
  uint32_t bin[kChunks];
  bool in[kChunks];
#pragma unroll
  for (uint32_t j = 0; j < kChunks; j++) {
    const uint32_t i = (c0 + j) * 64 + lane;
    in[j] = c0 + j < c1 && i < n;
    d[j] = in[j] ? a.key[tbase + t0 + i] : 0.0f;
    rec[j] = in[j] ? rec_in[t0 + i] : make_uint4(0, 0, 0, 0);
  }
  __syncthreads();
#pragma unroll
  for (uint32_t j = 0; j < kChunks; j++) {
    bin[j] = 0xFFFFu;                             // no survivor in this lane
    if (in[j]) {                                  // a NaN key (processOnlyVisible false) is a survivor too
      bin[j] = visit_bin(d[j], mn, rd, mode);
      atomicAdd(&s_off[wv][bin[j]], 1u);
    }
  }
  __syncthreads();
  // per bin: base of each wave (two bins per thread)
  if (P.ntiles > 1) {
    for (uint32_t j = 0; j < 2; j++) {
      const uint32_t k = tid + 256 * j;
      uint32_t run = col_run[j];
      for (uint32_t w = 0; w < kScatterWaves; w++) {
        const uint32_t c = s_off[w][k];
        s_off[w][k] = run;
        run += c;
      }
    }
  } else {
    uint32_t cnt[2], sum = 0;
    for (int j = 0; j < 2; j++) {
      cnt[j] = 0;
      for (uint32_t w = 0; w < kScatterWaves; w++) cnt[j] += s_off[w][2 * tid + j];
      sum += cnt[j];
    }
    uint32_t total;
    uint32_t run = block_exclusive_scan_256(sum, s_tmp, total);
    for (int j = 0; j < 2; j++) {
      const uint32_t k = 2 * tid + j;
      for (uint32_t w = 0; w < kScatterWaves; w++) {
        const uint32_t c = s_off[w][k];
        s_off[w][k] = run;
        run += c;
      }
    }
  }
  __syncthreads();

  // each wave ranks its own chunks; only this wave touches s_off[wv]
  const uint64_t lt_mask = (lane == 0) ? 0ull : (~0ull >> (64 - lane));
  uint32_t* off = s_off[wv];
#pragma unroll
  for (uint32_t j = 0; j < kChunks; j++) {
    const bool valid = bin[j] != 0xFFFFu;
    const uint64_t vmask = __ballot(valid);
    if (vmask == 0) continue;                     // uniform across the wave
    const uint64_t peers = bin_peers(bin[j], vmask);
    const uint32_t rank = (uint32_t)__popcll(peers & lt_mask);
    uint32_t slot = 0;
    if (valid) slot = off[bin[j]] + rank;
    __builtin_amdgcn_wave_barrier();
    if (valid && rank == 0) off[bin[j]] = slot + (uint32_t)__popcll(peers);
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    if (valid) rec_out[slot] = rec[j];
  }
}

typedef float f4v __attribute__((ext_vector_type(4)));

template <bool kNT>
__device__ __forceinline__ void store4(float4* p, float4 v) {
  if (kNT) {
    f4v x = {v.x, v.y, v.z, v.w};
    __builtin_nontemporal_store(x, reinterpret_cast<f4v*>(p));
  } else {
    *p = v;
  }
}
template <bool kNT>
__device__ __forceinline__ void store1(uint32_t* p, uint32_t v) {
  if (kNT) __builtin_nontemporal_store(v, p);
  else *p = v;
}

// Output-order assembly: per drawn triangle in slot order,
// the ebiten colour/depth vertices and indices (camera.go:1646-1820).  The
// sort pass left each slot's three vertex-record offsets in tri_vo, so the
// only gathers are the 32-byte vertex records and UVs.  kStage: stage each
// wavefront's 64 triangles in LDS and write contiguous float4 rows (else each
// lane stores its own 144-byte runs); kNT: non-temporal (streaming) stores.
// `split` blocks share one tile's slot range (a power of two dividing the
// tile), so the grid follows the survivors more finely than the sort's tiles.
// kOne: each slot range goes to two blocks, adjacent in launch order (so on
// one XCD, sharing the gathered records in its L2): the first writes the
// colour vertices and the indices, the second the depth vertices, so every
// block writes one contiguous output stream.
template <int kStage, bool kNT, uint32_t kWaves, bool kOne>
__global__ __launch_bounds__(64 * kWaves) void k_assemble(FrameArgs a, uint32_t blk0, uint32_t split) {
  constexpr uint32_t kHalves = kOne ? 2 : 1;
  const uint32_t lf = frame_block(a, FC_TILES, split * kHalves);
  if (lf == kNoBlock) return;
  const bool w_color = !kOne || (lf % kHalves) == 0;   // block-uniform
  const bool w_depth = !kOne || (lf % kHalves) == 1;
  const uint32_t li = lf / kHalves;
  const uint32_t b = blk0 + li / split;
  const uint32_t p = a.tile_part[b];
  const PartGPU P = a.parts[p];        // by value: read once, before any store
  const uint32_t S = a.stats[4 * p + 2];
  const uint32_t sub = a.tile / split;
  const uint32_t k0 = (b - a.tile_prefix[p]) * a.tile + (li % split) * sub;
  if (k0 >= S) return;                 // block-uniform: every wave of the block leaves together
  const uint32_t k1 = min(k0 + sub, S);
  const FrameConst fc = a.fc[P.cam];
  const uint32_t lane = threadIdx.x % 64, wave = threadIdx.x / 64;
  // Everything uniform is read once into registers here: the output stores
  // below may alias these tables as far as the compiler knows, and would
  // otherwise force a reload (a full memory round trip) per vertex.
  const uint32_t vbase = P.vbase, vmin = P.vmin, tbase = P.tbase;
  const bool has_uv = (P.flags & PF_HAS_UV) != 0;
  const bool unbb = (P.flags & PF_UNBB) != 0;
  const float* __restrict__ vud = a.vud;
  const float src_w = P.src_w, src_h = P.src_h;
  const float* __restrict__ uvp = P.uv;
  const uint64_t out_base = P.out_base;
  const uint32_t idx_base = P.idx_base;
  const bool persp = fc.perspective != 0, pc_tex = fc.pc_tex != 0, rdepth = fc.render_depth != 0;
  const float cam_w = fc.cam_w, half_w = fc.half_w, cam_h = fc.cam_h, half_h = fc.half_h;
  const float dmp = fc.dmp, spread = fc.spread;
  const float4* __restrict__ vclip = a.vclip;
  const float4* __restrict__ vcol = a.vcol;
  const uint32_t vstride = a.vstride;
  const uint4* __restrict__ tvo = a.tri_vo + tbase;
  float4* __restrict__ out_c = a.out_color + out_base * 3;   // 3 float4 per vertex
  float4* __restrict__ out_d = a.out_depth + out_base * 3;
  uint32_t* __restrict__ out_i = a.out_idx + out_base;
  const float da = rdepth ? 1.0f : 0.0f;

  // kStage 1: all 64 triangles of a wave staged at once; 2: 32 at a time,
  // halving the LDS per wavefront so twice as many wavefronts fit on a CU.
  // With kWaves > 1 the block's waves take consecutive 64-triangle chunks of
  // its slot range, each staging in its own LDS rows.
  constexpr uint32_t kRows = kStage == 2 ? 32 : 64;
  __shared__ float4 stage_all[kStage ? kWaves * kRows * 9 : 1];
  float4* stage = stage_all + (kStage ? wave * kRows * 9 : 0);
  constexpr uint32_t kStep = 64 * kWaves;
  uint4 vo_next = make_uint4(0, 0, 0, 0);   // slot records one chunk ahead
  if (k0 + wave * 64 + lane < k1) vo_next = tvo[k0 + wave * 64 + lane];
  for (uint32_t base = k0; base < k1; base += kStep) {   // block-uniform trip count
    const uint32_t wk0 = base + wave * 64;
    const uint32_t k = wk0 + lane;
    const bool active = k < k1;
    const uint32_t wcount = wk0 < k1 ? min(64u, k1 - wk0) : 0u;
    const uint4 vo = vo_next;
    if (k + kStep < k1) vo_next = tvo[k + kStep];
    float4 cq[9];
    float dd[3];
    if (active) {
      const uint32_t ov[3] = {vo.x, vo.y, vo.z};
      // all nine gathers of the triangle first, so their latencies overlap
      float4 c[3], col[3];
      float2 uv[3];
      for (int vi = 0; vi < 3; vi++) {
        c[vi] = vclip[(size_t)ov[vi] * vstride];
        col[vi] = w_color ? vcol[(size_t)ov[vi] * vstride] : make_float4(0.0f, 0.0f, 0.0f, 0.0f);
        uv[vi] = has_uv ? reinterpret_cast<const float2*>(uvp)[ov[vi] - vbase + vmin] : make_float2(0.0f, 0.0f);
      }
      for (int vi = 0; vi < 3; vi++) {
        float w = c[vi].w;                                        // camera.go:1655-1667
        if (!persp) w = 1.0f;
        if (w < 0) w = 0.001f;
        const float dx = fdiv(c[vi].x, w) * cam_w + half_w;
        const float dy = fdiv(c[vi].y, -w) * cam_h + half_h;
        const float U = uv[vi].x, V = uv[vi].y;
        float uvU, uvV, c0 = 0.0f;
        if (pc_tex) {                                             // camera.go:1688-1711
          uvU = fdiv(U, w) * src_w;
          uvV = fdiv(1.0f - V, w) * src_h;
          c0 = fdiv(1.0f, w);
        } else {
          uvU = U * src_w;
          uvV = (1.0f - V) * src_h;
        }
        float depth = 0.0f;
        if (rdepth && w_depth) {                                  // camera.go:1765-1784
          const float zd = unbb ? vud[ov[vi]] : c[vi].z;          // DepthModeUnbillboarded: the unaltered Z
          depth = fdiv(zd + dmp, spread);
          if (depth < 0) depth = 0;
          else if (depth > 1) depth = 1;
        }
        cq[3 * vi + 0] = make_float4(dx, dy, uvU, uvV);
        cq[3 * vi + 1] = col[vi];
        cq[3 * vi + 2] = make_float4(c0, 0.0f, 0.0f, 0.0f);
        dd[vi] = depth;
      }
    }
    if (kStage) {
      for (uint32_t h = 0; h < 64 / kRows; h++) {
        const uint32_t r0 = h * kRows;                              // first row of this round
        const uint32_t rows = wcount > r0 ? min(kRows, wcount - r0) : 0u;
        const bool mine = active && lane >= r0 && lane < r0 + kRows;
        const uint32_t row = lane - r0;
        if (w_color) {
          if (mine)
            for (int q = 0; q < 9; q++) stage[row * 9 + q] = cq[q];
          __syncthreads();
          // a fixed number of (predicated) stores per lane, so the compiler
          // can count them and wait only for the loads it needs
#pragma unroll
          for (uint32_t q = 0; q < (kRows * 9 + 63) / 64; q++) {
            const uint32_t f = q * 64 + lane;
            if (f < rows * 9) store4<kNT>(&out_c[(size_t)(wk0 + r0) * 9 + f], stage[f]);
          }
          __syncthreads();
        }
        if (w_depth) {
          if (mine)
            for (int vi = 0; vi < 3; vi++) {
              if (!w_color) {                                     // rows not staged for the colour list
                stage[row * 9 + 3 * vi + 0] = cq[3 * vi + 0];
                stage[row * 9 + 3 * vi + 2] = cq[3 * vi + 2];
              }
              stage[row * 9 + 3 * vi + 1] = make_float4(dd[vi], dd[vi], dd[vi], da);
            }
          __syncthreads();
#pragma unroll
          for (uint32_t q = 0; q < (kRows * 9 + 63) / 64; q++) {
            const uint32_t f = q * 64 + lane;
            if (f < rows * 9) store4<kNT>(&out_d[(size_t)(wk0 + r0) * 9 + f], stage[f]);
          }
          __syncthreads();
        }
      }
    } else if (active) {
      float4* oc = out_c + (size_t)k * 9;
      float4* od = out_d + (size_t)k * 9;
      if (w_color)
        for (int q = 0; q < 9; q++) store4<kNT>(&oc[q], cq[q]);
      if (w_depth)
        for (int vi = 0; vi < 3; vi++) {
          store4<kNT>(&od[3 * vi + 0], cq[3 * vi + 0]);
          store4<kNT>(&od[3 * vi + 1], make_float4(dd[vi], dd[vi], dd[vi], da));
          store4<kNT>(&od[3 * vi + 2], cq[3 * vi + 2]);
        }
    }
    // indices: flush-local iota, as the reference's indexList (camera.go:1816)
    if (w_color) {
#pragma unroll
      for (uint32_t q = 0; q < 3; q++) {
        const uint32_t f = q * 64 + lane;
        if (f < wcount * 3) store1<kNT>(&out_i[(size_t)wk0 * 3 + f], idx_base + wk0 * 3 + f);
      }
    }
  }
}

// The same assembly, software-pipelined: the nine gathers of chunk i+1 are
// issued before chunk i's stores, so their latency hides behind the stores.
// Every memory instruction runs on all 64 lanes (lanes past the end read a
// clamped slot and write to a sink), which keeps the number of stores per
// chunk fixed: the wait for the next chunk's gathers can then let the
// current chunk's stores stay in flight instead of draining them.
template <uint32_t kRows, bool kNT>
__global__ __launch_bounds__(64) void k_assemble_pipe(FrameArgs a, uint32_t blk0, uint32_t split) {
  const uint32_t li = frame_block(a, FC_TILES, split);
  if (li == kNoBlock) return;
  const uint32_t b = blk0 + li / split;
  const uint32_t p = a.tile_part[b];
  const PartGPU P = a.parts[p];
  const uint32_t S = a.stats[4 * p + 2];
  const uint32_t sub = a.tile / split;
  const uint32_t k0 = (b - a.tile_prefix[p]) * a.tile + (li % split) * sub;
  if (k0 >= S) return;
  const uint32_t k1 = min(k0 + sub, S), last = k1 - 1;
  const FrameConst fc = a.fc[P.cam];
  const uint32_t lane = threadIdx.x;
  const uint32_t vbase = P.vbase, vmin = P.vmin, tbase = P.tbase;
  const bool has_uv = (P.flags & PF_HAS_UV) != 0;
  const bool unbb = (P.flags & PF_UNBB) != 0;
  const float* __restrict__ vud = a.vud;
  const float src_w = P.src_w, src_h = P.src_h;
  const float2* __restrict__ uvp = reinterpret_cast<const float2*>(P.uv);
  const uint64_t out_base = P.out_base;
  const uint32_t idx_base = P.idx_base;
  const bool persp = fc.perspective != 0, pc_tex = fc.pc_tex != 0, rdepth = fc.render_depth != 0;
  const float cam_w = fc.cam_w, half_w = fc.half_w, cam_h = fc.cam_h, half_h = fc.half_h;
  const float dmp = fc.dmp, spread = fc.spread;
  const float4* __restrict__ vclip = a.vclip;
  const float4* __restrict__ vcol = a.vcol;
  const uint32_t vstride = a.vstride;
  const uint4* __restrict__ tvo = a.tri_vo + tbase;
  float4* __restrict__ out_c = a.out_color + out_base * 3;
  float4* __restrict__ out_d = a.out_depth + out_base * 3;
  uint32_t* __restrict__ out_i = a.out_idx + out_base;
  float4* __restrict__ sink = a.sink + lane;                 // per-lane sink, never read
  uint32_t* __restrict__ sink_i = reinterpret_cast<uint32_t*>(a.sink + 64) + lane;
  const float da = rdepth ? 1.0f : 0.0f;
  __shared__ float4 stage[kRows * 9];

  struct Gathered {
    float4 c[3], col[3];
    float2 uv[3];
    float zd[3];                 // Z the depth list takes: clip Z, or the unbillboarded Z
  };
  auto gather = [&](Gathered& g, const uint4 vo) {
    const uint32_t ov[3] = {vo.x, vo.y, vo.z};
    for (int vi = 0; vi < 3; vi++) {
      g.c[vi] = vclip[(size_t)ov[vi] * vstride];
      g.col[vi] = vcol[(size_t)ov[vi] * vstride];
      g.uv[vi] = has_uv ? uvp[ov[vi] - vbase + vmin] : make_float2(0.0f, 0.0f);
      g.zd[vi] = unbb ? vud[ov[vi]] : g.c[vi].z;
    }
  };
  auto compute = [&](const Gathered& g, float4 (&cq)[9], float (&dd)[3]) {
    for (int vi = 0; vi < 3; vi++) {
      float w = g.c[vi].w;                                        // camera.go:1655-1667
      if (!persp) w = 1.0f;
      if (w < 0) w = 0.001f;
      const float dx = fdiv(g.c[vi].x, w) * cam_w + half_w;
      const float dy = fdiv(g.c[vi].y, -w) * cam_h + half_h;
      const float U = g.uv[vi].x, V = g.uv[vi].y;
      float uvU, uvV, c0 = 0.0f;
      if (pc_tex) {                                               // camera.go:1688-1711
        uvU = fdiv(U, w) * src_w;
        uvV = fdiv(1.0f - V, w) * src_h;
        c0 = fdiv(1.0f, w);
      } else {
        uvU = U * src_w;
        uvV = (1.0f - V) * src_h;
      }
      float depth = 0.0f;
      if (rdepth) {                                               // camera.go:1765-1784
        depth = fdiv(g.zd[vi] + dmp, spread);
        if (depth < 0) depth = 0;
        else if (depth > 1) depth = 1;
      }
      cq[3 * vi + 0] = make_float4(dx, dy, uvU, uvV);
      cq[3 * vi + 1] = g.col[vi];
      cq[3 * vi + 2] = make_float4(c0, 0.0f, 0.0f, 0.0f);
      dd[vi] = depth;
    }
  };
  auto store_chunk = [&](const float4 (&cq)[9], const float (&dd)[3], uint32_t wk0) {
    const uint32_t wcount = min(64u, k1 - wk0);
#pragma unroll
    for (uint32_t h = 0; h < 64 / kRows; h++) {
      const uint32_t r0 = h * kRows;
      const uint32_t rows = wcount > r0 ? min(kRows, wcount - r0) : 0u;
      const bool mine = lane >= r0 && lane < r0 + kRows;
      if (mine)
        for (int q = 0; q < 9; q++) stage[(lane - r0) * 9 + q] = cq[q];
      __syncthreads();
#pragma unroll
      for (uint32_t q = 0; q < (kRows * 9 + 63) / 64; q++) {
        const uint32_t f = q * 64 + lane;
        store4<kNT>(f < rows * 9 ? &out_c[(size_t)(wk0 + r0) * 9 + f] : sink, stage[min(f, kRows * 9 - 1)]);
      }
      __syncthreads();
      if (mine)
        for (int vi = 0; vi < 3; vi++) stage[(lane - r0) * 9 + 3 * vi + 1] = make_float4(dd[vi], dd[vi], dd[vi], da);
      __syncthreads();
#pragma unroll
      for (uint32_t q = 0; q < (kRows * 9 + 63) / 64; q++) {
        const uint32_t f = q * 64 + lane;
        store4<kNT>(f < rows * 9 ? &out_d[(size_t)(wk0 + r0) * 9 + f] : sink, stage[min(f, kRows * 9 - 1)]);
      }
      __syncthreads();
    }
#pragma unroll
    for (uint32_t q = 0; q < 3; q++) {                            // indices (camera.go:1816)
      const uint32_t f = q * 64 + lane;
      store1<kNT>(f < wcount * 3 ? &out_i[(size_t)wk0 * 3 + f] : sink_i, idx_base + wk0 * 3 + f);
    }
  };

  // two register sets, used in turn: no copies between chunks
  Gathered ga, gb;
  gather(ga, tvo[min(k0 + lane, last)]);
  uint4 vo_next = tvo[min(k0 + 64 + lane, last)];
  float4 cq[9];
  float dd[3];
  for (uint32_t wk0 = k0;;) {
    compute(ga, cq, dd);
    bool more = wk0 + 64 < k1;                                    // uniform
    if (more) {
      gather(gb, vo_next);
      vo_next = tvo[min(wk0 + 128 + lane, last)];
    }
    store_chunk(cq, dd, wk0);
    wk0 += 64;
    if (!more) break;
    compute(gb, cq, dd);
    more = wk0 + 64 < k1;
    if (more) {
      gather(ga, vo_next);
      vo_next = tvo[min(wk0 + 128 + lane, last)];
    }
    store_chunk(cq, dd, wk0);
    wk0 += 64;
    if (!more) break;
  }
}

// Fill of a word array (test poisoning of outputs: unwritten entries show up).
__global__ __launch_bounds__(256) void k_fill(uint32_t* __restrict__ p, uint64_t n, uint32_t v) {
  for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x) p[i] = v;
}

// Survivors of a range of pairs packed back to back for the download (one
// transfer per stream instead of three per pair).  A block finds its pair by
// binary search of the block prefix: the pair p with blk_prefix[p] <= block <
// blk_prefix[p + 1] (every listed pair has at least one vertex, and 3 * nvert
// elements, 256 to a block).
__global__ __launch_bounds__(256) void k_pack(PackArgs a) {
  const uint32_t b = blockIdx.x;
  uint32_t lo = 0, hi = a.npairs;
  while (hi - lo > 1) {
    const uint32_t mid = (lo + hi) >> 1;
    if (a.blk_prefix[mid] <= b) lo = mid;
    else hi = mid;
  }
  // element e of the pair: float4 e of its colour and depth runs (3 per vertex), and index e while e < nvert,
  // so consecutive lanes move consecutive 16-byte words
  const uint64_t e = (uint64_t)(b - a.blk_prefix[lo]) * 256 + threadIdx.x;
  const uint64_t n = a.nvert[lo];
  if (e >= 3 * n) return;
  const uint64_t s = a.src[lo], d = a.dst[lo];
  if (a.pcolor) a.pcolor[3 * d + e] = a.color[3 * s + e];
  if (a.pdepth) a.pdepth[3 * d + e] = a.depth[3 * s + e];
  if (a.pnormal) a.pnormal[3 * d + e] = a.normal[3 * s + e];
  if (a.pidx && e < n) a.pidx[d + e] = a.idx[s + e];
}

// Dynamic batches (camera.go:2075-2099): the render pairs of one flush write
// into one vertex list, each after the corners of the ones before it, and
// their indices count on from there (the reference's vertexListIndex).  After
// k_cull, one block per batch scans its pairs' survivor counts and moves each
// pair's out_base and idx_base accordingly.  Only pairs in [part0, part1) are
// written (the plain form runs one pair's chain at a time; the pairs before
// it in its flush are final by then).  The first pair of a flush keeps the
// flush's own out_base, which every block reads before any write.
__global__ __launch_bounds__(256) void k_groups(FrameArgs a, uint32_t part0, uint32_t part1) {
  const uint2 g = a.groups[blockIdx.x];
  const uint32_t first = g.x, end = g.x + g.y;
  if (end <= part0 || first >= part1) return;                       // block-uniform
  PartGPU* parts = const_cast<PartGPU*>(a.parts);
  const uint64_t base = parts[first].out_base;
  __shared__ uint32_t s_scan[256];
  __shared__ uint32_t s_carry;
  if (threadIdx.x == 0) s_carry = 0;
  __syncthreads();
  for (uint32_t c0 = first; c0 < end; c0 += 256) {                  // block-uniform trip count
    const uint32_t i = c0 + threadIdx.x;
    const uint32_t v = i < end ? a.stats[4 * i + 2] : 0u;
    s_scan[threadIdx.x] = v;
    __syncthreads();
    for (uint32_t d = 1; d < 256; d <<= 1) {                        // inclusive scan
      const uint32_t t = threadIdx.x >= d ? s_scan[threadIdx.x - d] : 0u;
      __syncthreads();
      s_scan[threadIdx.x] += t;
      __syncthreads();
    }
    const uint32_t before = s_carry + s_scan[threadIdx.x] - v;      // survivors of the flush's earlier pairs
    if (i < end && i >= part0 && i < part1) {
      parts[i].out_base = base + 3ull * before;
      parts[i].idx_base = 3u * before;
    }
    __syncthreads();
    if (threadIdx.x == 255) s_carry += s_scan[255];
    __syncthreads();
  }
}

// Camera.RenderNormals: the normal vertex list (camera.go:1710-1722), after
// the assembly.  Per drawn triangle in slot order, each corner takes the
// colour vertex's DstX, DstY, SrcX, SrcY (its first float4, just written)
// and as colour the vertex's rotated position times 0.5 plus 0.5
// (k_vertex's vnrm); alpha and the custom fields stay 0.  One 256-thread
// block per sort tile, one thread per (triangle, corner).
__global__ __launch_bounds__(256) void k_normals(FrameArgs a, uint32_t blk0) {
  const uint32_t lb = frame_block(a, FC_TILES, 1);
  if (lb == kNoBlock) return;
  const uint32_t b = blk0 + lb;
  const uint32_t p = a.tile_part[b];
  const PartGPU P = a.parts[p];
  if (!a.fc[P.cam].render_normals) return;
  const uint32_t S = a.stats[4 * p + 2];
  const uint32_t k0 = (b - a.tile_prefix[p]) * a.tile;
  if (k0 >= S) return;
  const uint32_t k1 = min(k0 + a.tile, S);
  const uint4* __restrict__ tvo = a.tri_vo + P.tbase;
  const float4* __restrict__ oc = a.out_color + P.out_base * 3;
  float4* __restrict__ on = a.out_normal + P.out_base * 3;
  const float4* __restrict__ vn = a.vnrm;
  for (uint32_t c = 3 * k0 + threadIdx.x; c < 3 * k1; c += blockDim.x) {   // corner c of the pair's list
    const uint4 vo = tvo[c / 3];
    const uint32_t vi = c % 3;
    const uint32_t o = vi == 0 ? vo.x : (vi == 1 ? vo.y : vo.z);
    on[3 * (size_t)c + 0] = oc[3 * (size_t)c + 0];
    on[3 * (size_t)c + 1] = vn[o];
    on[3 * (size_t)c + 2] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
  }
}

// Plain float4 copy of a known size: the reference point for HBM byte counters.
__global__ __launch_bounds__(256) void k_copy(const float4* __restrict__ src, float4* __restrict__ dst, uint64_t n) {
  for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x)
    dst[i] = src[i];
}

__global__ void k_fp_selftest(const float* x, const float* y, float* out, uint32_t n) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  out[i] = fdiv(x[i], y[i]);
  out[n + i] = fsqrt(fabsf(x[i]));
  out[2 * n + i] = x[i] * y[i] + x[i];
  out[3 * n + i] = roundf(x[i]);
}

// ---------------------------------------------------------------------------
// launchers
// ---------------------------------------------------------------------------

hipError_t launch_palettes(const FrameArgs& a, hipStream_t s) {
  if (!a.npal_jobs) return hipSuccess;
  hipLaunchKernelGGL(k_palettes, dim3(a.npal_jobs), dim3(kPoseThreads), 0, s, a);
  return hipGetLastError();
}
hipError_t launch_uniforms(const FrameArgs& a, uint32_t part0, uint32_t nparts, hipStream_t s) {
  if (!nparts) return hipSuccess;
  hipLaunchKernelGGL(k_uniforms, dim3(nparts), dim3(64), 0, s, a, part0, nparts);
  return hipGetLastError();
}
hipError_t launch_vertex(const FrameArgs& a, uint32_t blk0, uint32_t nblk, uint32_t threads, hipStream_t s) {
  if (!nblk) return hipSuccess;
  hipLaunchKernelGGL(k_vertex, dim3(nblk), dim3(threads), 0, s, a, blk0);
  return hipGetLastError();
}
hipError_t launch_cull(const FrameArgs& a, uint32_t blk0, uint32_t nblk, uint32_t threads, hipStream_t s) {
  if (!nblk) return hipSuccess;
  switch (threads) {
    case 256: hipLaunchKernelGGL(k_cull<256>, dim3(nblk), dim3(256), 0, s, a, blk0); break;
    case 512: hipLaunchKernelGGL(k_cull<512>, dim3(nblk), dim3(512), 0, s, a, blk0); break;
    case 1024: hipLaunchKernelGGL(k_cull<1024>, dim3(nblk), dim3(1024), 0, s, a, blk0); break;
    default: return hipErrorInvalidValue;
  }
  return hipGetLastError();
}
hipError_t launch_hist(const FrameArgs& a, uint32_t blk0, uint32_t nblk, hipStream_t s) {
  if (!nblk) return hipSuccess;
  hipLaunchKernelGGL(k_hist, dim3(nblk), dim3(kHistThreads), 0, s, a, blk0);
  return hipGetLastError();
}
hipError_t launch_scan(const FrameArgs& a, uint32_t sblk0, uint32_t nsblk, uint32_t part0, uint32_t nparts,
                       hipStream_t s) {
  if (!nsblk) return hipSuccess;
  hipLaunchKernelGGL(k_scan_reduce, dim3(nsblk), dim3(256), 0, s, a, sblk0);
  hipLaunchKernelGGL(k_scan_bsums, dim3(nparts), dim3(256), 0, s, a, part0);
  hipLaunchKernelGGL(k_scan_down, dim3(nsblk), dim3(256), 0, s, a, sblk0);
  return hipGetLastError();
}
hipError_t launch_scan_part(const FrameArgs& a, uint32_t blk0, uint32_t nblk, hipStream_t s) {
  if (!nblk) return hipSuccess;
  hipLaunchKernelGGL(k_scan_part, dim3(nblk), dim3(1024), 0, s, a, blk0);
  return hipGetLastError();
}
hipError_t launch_scatter(const FrameArgs& a, uint32_t blk0, uint32_t nblk, uint32_t waves, hipStream_t s) {
  if (!nblk) return hipSuccess;
  if (waves == kScatterWaves) hipLaunchKernelGGL(k_scatter4, dim3(nblk), dim3(64 * kScatterWaves), 0, s, a, blk0);
  else hipLaunchKernelGGL(k_scatter, dim3(nblk), dim3(64), 0, s, a, blk0);
  return hipGetLastError();
}
hipError_t launch_assemble(const FrameArgs& a, uint32_t blk0, uint32_t nblk, uint32_t mode, uint32_t split,
                           hipStream_t s) {
  if (!nblk) return hipSuccess;
  if (split == 0 || (split & (split - 1)) || a.tile % split) return hipErrorInvalidValue;
  if (mode & 64u) {   // bit 6: the assembly's blocks in launch order, whatever the other passes use
    FrameArgs b = a;
    b.xcd_swizzle = 0;
    return launch_assemble(b, blk0, nblk, mode & ~64u, split, s);
  }
  const dim3 g(nblk * split), g2(2 * nblk * split);
  switch (mode) {
    case 8: hipLaunchKernelGGL((k_assemble_pipe<64, false>), g, dim3(64), 0, s, a, blk0, split); break;
    case 10: hipLaunchKernelGGL((k_assemble_pipe<64, true>), g, dim3(64), 0, s, a, blk0, split); break;
    case 12: hipLaunchKernelGGL((k_assemble_pipe<32, false>), g, dim3(64), 0, s, a, blk0, split); break;
    case 14: hipLaunchKernelGGL((k_assemble_pipe<32, true>), g, dim3(64), 0, s, a, blk0, split); break;
    case 0: hipLaunchKernelGGL((k_assemble<1, false, 1, false>), g, dim3(64), 0, s, a, blk0, split); break;
    case 1: hipLaunchKernelGGL((k_assemble<0, false, 1, false>), g, dim3(64), 0, s, a, blk0, split); break;
    case 2: hipLaunchKernelGGL((k_assemble<1, true, 1, false>), g, dim3(64), 0, s, a, blk0, split); break;
    case 3: hipLaunchKernelGGL((k_assemble<0, true, 1, false>), g, dim3(64), 0, s, a, blk0, split); break;
    case 4: hipLaunchKernelGGL((k_assemble<2, false, 1, false>), g, dim3(64), 0, s, a, blk0, split); break;
    case 6: hipLaunchKernelGGL((k_assemble<2, true, 1, false>), g, dim3(64), 0, s, a, blk0, split); break;
    // bit 4: four wavefronts per block, each staging its own rows
    case 16: hipLaunchKernelGGL((k_assemble<1, false, 4, false>), g, dim3(256), 0, s, a, blk0, split); break;
    case 18: hipLaunchKernelGGL((k_assemble<1, true, 4, false>), g, dim3(256), 0, s, a, blk0, split); break;
    case 20: hipLaunchKernelGGL((k_assemble<2, false, 4, false>), g, dim3(256), 0, s, a, blk0, split); break;
    case 22: hipLaunchKernelGGL((k_assemble<2, true, 4, false>), g, dim3(256), 0, s, a, blk0, split); break;
    // bit 5: one output stream per block (colour vertices and indices, or depth vertices), twice the blocks
    case 34: hipLaunchKernelGGL((k_assemble<1, true, 1, true>), g2, dim3(64), 0, s, a, blk0, split); break;
    case 38: hipLaunchKernelGGL((k_assemble<2, true, 1, true>), g2, dim3(64), 0, s, a, blk0, split); break;
    case 50: hipLaunchKernelGGL((k_assemble<1, true, 4, true>), g2, dim3(256), 0, s, a, blk0, split); break;
    case 54: hipLaunchKernelGGL((k_assemble<2, true, 4, true>), g2, dim3(256), 0, s, a, blk0, split); break;
    default: return hipErrorInvalidValue;
  }
  return hipGetLastError();
}
hipError_t launch_fill(void* dst, uint64_t bytes, uint32_t word, hipStream_t s) {
  if (bytes < 4) return hipSuccess;
  hipLaunchKernelGGL(k_fill, dim3(2048), dim3(256), 0, s, reinterpret_cast<uint32_t*>(dst), bytes / 4, word);
  return hipGetLastError();
}
hipError_t launch_pack(const PackArgs& a, uint32_t nblocks, hipStream_t s) {
  if (!nblocks) return hipSuccess;
  hipLaunchKernelGGL(k_pack, dim3(nblocks), dim3(256), 0, s, a);
  return hipGetLastError();
}
hipError_t launch_normals(const FrameArgs& a, uint32_t blk0, uint32_t nblk, hipStream_t s) {
  if (!nblk) return hipSuccess;
  hipLaunchKernelGGL(k_normals, dim3(nblk), dim3(256), 0, s, a, blk0);
  return hipGetLastError();
}
hipError_t launch_groups(const FrameArgs& a, uint32_t part0, uint32_t part1, hipStream_t s) {
  if (!a.ngroups) return hipSuccess;
  hipLaunchKernelGGL(k_groups, dim3(a.ngroups), dim3(256), 0, s, a, part0, part1);
  return hipGetLastError();
}
hipError_t launch_copy(const void* src, void* dst, uint64_t bytes, hipStream_t s) {
  const uint64_t n = bytes / 16;
  hipLaunchKernelGGL(k_copy, dim3(2048), dim3(256), 0, s, reinterpret_cast<const float4*>(src), reinterpret_cast<float4*>(dst), n);
  return hipGetLastError();
}
hipError_t launch_fp_selftest(const float* x, const float* y, float* out, uint32_t n, hipStream_t s) {
  hipLaunchKernelGGL(k_fp_selftest, dim3((n + 255) / 256), dim3(256), 0, s, x, y, out, n);
  return hipGetLastError();
}

}  // namespace t3d
