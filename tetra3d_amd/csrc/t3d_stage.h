// t3d_stage.h — device functions of the model-level stage shared by
// k_models/k_plist (t3d_models.hip) and k_uniforms (t3d_kernels.hip), which
// builds each render pair's PartGPU of a device-built frame.
#pragma once

#include "t3d_internal.h"

namespace t3d {

// part_material (a material array whose entry is present), or null
__device__ __forceinline__ const MaterialGPU* stage_material(const ModelStage& ms, const ModelGPU& M, uint32_t part) {
  if (!(M.flags & MF_MATERIALS)) return nullptr;
  const MaterialGPU* mat = ms.mats + M.mat + part;
  return (mat->bits & MB_PRESENT) ? mat : nullptr;
}

// Model.isTransparent (model.go:1274-1282)
__device__ __forceinline__ bool stage_transparent(const float* mcolor, const MaterialGPU* mat) {
  if (mat) {
    const bool matT = mat->transparency_mode == T3D_TRANSPARENCY_TRANSPARENT || !(mat->bits & MB_BLEND_SOURCE_OVER) ||
                      (mat->transparency_mode == T3D_TRANSPARENCY_AUTO && mat->color[3] < 0.999f);
    const bool modelT = mat->transparency_mode != T3D_TRANSPARENCY_OPAQUE && mcolor[3] < 0.999f;
    return matT || modelT;
  }
  return mcolor[3] < 0.999f;
}

// What one render pair of (model, part) draws and the slices it needs
// (camera.go:1521-1528 lighting, model.go:604-606 early out).
struct PairSizes {
  uint32_t vcount, tcount, nvchunks, ntiles;
  uint32_t nsblocks, pscan, hist, lcount;
  uint32_t bytes_vertex, bytes_cull, bytes_uv;
  bool lighting;
};
__device__ __forceinline__ PairSizes pair_sizes(const ModelStage& ms, const ModelGPU& M, const MeshGPU& mesh,
                                                const MeshPartGPU& pr, const MaterialGPU* mat, bool transparent) {
  PairSizes z{};
  const JobGPU& J = ms.jobs[M.cam];
  z.lighting = J.lighting_on && !(M.flags & MF_SHADELESS) && !(mat && (mat->bits & MB_SHADELESS));
  const bool early_out = M.color[3] == 0.0f && transparent;
  const bool skinned = (M.flags & MF_SKINNED) != 0;
  z.vcount = early_out ? 0u : pr.vcount;
  z.tcount = early_out ? 0u : pr.triangle_count;
  z.nvchunks = (z.vcount + ms.vchunk - 1) / ms.vchunk;
  z.ntiles = (z.tcount + ms.tile - 1) / ms.tile;
  if (z.ntiles > 1) {
    z.hist = kBins * z.ntiles;
    if (z.hist <= kScanPartMax) z.pscan = 1;   // scanned whole by one k_scan_part block
    else z.nsblocks = (z.hist + kScanBlock - 1) / kScanBlock;
  }
  z.lcount = z.lighting ? M.light_count : 0u;
  const uint32_t vb = 24u + (mesh.col ? 16u : 0u) + (skinned ? 24u : 0u);
  z.bytes_vertex = z.vcount * vb + (skinned ? 64u * M.bone_count : 0u) + (uint32_t)sizeof(PartGPU);
  z.bytes_cull = ((mesh.idx16 ? 6u : 12u) + (skinned ? 0u : 12u)) * z.tcount;
  z.bytes_uv = mesh.uv ? 8u * z.vcount : 0u;
  return z;
}

// The PartGPU of render pair p of a device-built frame, as the host's
// emit_pair writes it (camera.go:1507-1606, model.go:600-695): mesh arrays,
// flags, colour, lights, the billboarded base; k_uniforms completes it.
__device__ __forceinline__ void build_pair(const FrameArgs& a, uint32_t p, PartGPU& P) {
  const ModelStage& ms = a.ms;
  const uint32_t c = ms.pair_cand[p];
  const ModelGPU& M = ms.models[ms.cand_model[c]];
  const uint32_t part = c - M.cand0;
  const MeshGPU& mesh = *M.mesh;
  const MeshPartGPU pr = mesh.parts[part];
  const MaterialGPU* mat = stage_material(ms, M, part);
  const JobGPU& J = ms.jobs[M.cam];
  const bool skinned = (M.flags & MF_SKINNED) != 0;
  const bool transparent = stage_transparent(M.color, mat);
  const PairSizes z = pair_sizes(ms, M, mesh, pr, mat, transparent);
  P = PartGPU{};
  P.model = M.transform;
  if (mat && (mat->bits & MB_BILLBOARD) && !skinned)   // model.go:634-693
    P.mvp = billboard_base(M.transform, v3(J.pos[0], J.pos[1], J.pos[2]),
                           (mat->bits & MB_UP_CAMERA_Y) ? v3(J.up[0], J.up[1], J.up[2]) : v3(0.0f, 1.0f, 0.0f),
                           (mat->bits & MB_LOCK_X) != 0, (mat->bits & MB_LOCK_Y) != 0, (mat->bits & MB_LOCK_Z) != 0);
  else
    P.mvp = M.transform;
  P.cam = M.cam;
  P.max_span = mesh.max_span;
  for (int i = 0; i < 4; i++) P.mp[i] = mat ? M.color[i] * mat->color[i] : M.color[i];   // camera.go:1597-1601
  if (mat && (mat->bits & MB_TEXTURE)) {
    P.src_w = mat->texture_size[0];
    P.src_h = mat->texture_size[1];
  }
  uint32_t flags = 0;
  if (skinned) flags |= PF_SKINNED;
  if (mat && (mat->bits & MB_BACKFACE)) flags |= PF_BACKFACE;
  if (z.lighting) flags |= PF_LIGHTING;
  if (mesh.col) flags |= PF_HAS_COLOR;
  if (mesh.uv) flags |= PF_HAS_UV;
  if (mesh.K) flags |= PF_SHAPE_KEYS;
  if (mat && (mat->bits & MB_UNBILLBOARDED)) flags |= PF_UNBB;
  flags |= (mat ? (mat->modes & 3u) : (uint32_t)T3D_SORT_BACK_TO_FRONT) << PF_SORT_SHIFT;
  flags |= (mat ? ((mat->modes >> 2) & 3u) : 0u) << PF_LMODE_SHIFT;
  flags |= (mat ? ((mat->modes >> 4) & 3u) : 0u) << PF_LVMODE_SHIFT;
  const size_t mp = ms.max_pairs;
  if (z.lighting) {
    P.light_base = ms.pc[7 * mp + p];
    P.light_count = M.light_count;
    P.light0 = M.light0;
    if ((M.flags & MF_DIRECTIONAL) && !skinned) {   // k_uniforms inverts it (light.go:326-330)
      P.rd = M.world_rotation;
      flags |= PF_NEED_RD;
    }
  }
  P.flags = flags;
  P.pos = mesh.pos;
  P.nrm = mesh.nrm;
  P.col = mesh.col;
  P.uv = mesh.uv;
  P.bones = mesh.bones;
  P.weights = mesh.weights;
  P.idx = mesh.idx ? mesh.idx + 3 * (size_t)pr.triangle_start : nullptr;
  P.idx16 = mesh.idx16 ? mesh.idx16 + 3 * (size_t)pr.triangle_start : nullptr;
  P.ctr = mesh.ctr + 3 * (size_t)pr.triangle_start;
  P.sk_pos = mesh.sk_pos;
  P.sk_nrm = mesh.sk_nrm;
  P.shape_keys = mesh.K;
  P.mesh_vcount = mesh.V;
  P.bone_count = skinned ? M.bone_count : 0u;
  P.palette = skinned ? ms.pal_base + M.palette : nullptr;
  P.sk_w = mesh.K ? ms.skw_base + M.skw : nullptr;
  P.vmin = pr.vmin;
  P.vcount = z.vcount;
  P.tcount = z.tcount;
  P.vbase = ms.pc[0 * mp + p];
  P.tbase = ms.pc[1 * mp + p];
  P.out_base = 3ull * P.tbase;
  P.ntiles = z.ntiles;
  if (z.ntiles > 1) {
    P.hbase = ms.pc[6 * mp + p];
    P.sblock0 = ms.sblock_prefix[p];
    P.nsblocks = z.nsblocks;
  }
}

}  // namespace t3d
