// t3d_models.hip — the model-level stage of Camera.Render on the device: the
// frustum test of every model, the render-pair list and its order, and the
// sizes of every pair's slices of the frame arrays.
//
//  k_models  one thread per drawable model: its frustum sphere
//            (Model.onTransformUpdate, model.go:142-182) against its camera
//            (SphereInFrustum, camera.go:933-983), then per MeshPart one
//            candidate: drawn or not (frustum, Material.Visible), solid or
//            transparent (model.go:1274-1282), what it would draw.
//  k_plist   one block: the render pairs of each Camera.Render in the
//            reference's order (camera.go:1343-1457, 2050-2109): its solid
//            candidates in model order, then its transparents sorted by
//            (RenderOrder ascending, distance² descending), stable; then one
//            scan over the pairs gives every slice (vertex records, triangle
//            scratch, histograms, scan blocks, per-light values, outputs) and
//            the prefix arrays the chain's kernels find their pair by.
// k_uniforms (t3d_kernels.hip) builds each pair's PartGPU from the model,
// mesh and material records (build_pair below).
#include <hip/hip_runtime.h>

#include "t3d_internal.h"
#include "t3d_stage.h"

namespace t3d {

namespace {

// ---------------------------------------------------------------------------
// k_models
// ---------------------------------------------------------------------------

constexpr uint32_t kModelThreads = 64;   // one wavefront per block: the per-model work is serial ALU, spread it over CUs

__global__ __launch_bounds__(kModelThreads) void k_models(FrameArgs a) {
  const ModelStage& ms = a.ms;
  const uint32_t m = blockIdx.x * kModelThreads + threadIdx.x;
  if (m >= ms.nmodels) return;
  const ModelGPU& M = ms.models[m];
  const MeshGPU& mesh = *M.mesh;
  const JobGPU& J = ms.jobs[M.cam];
  const bool skinned = (M.flags & MF_SKINNED) != 0;
  bool in = true;
  if (M.flags & MF_FRUSTUM) {
    // Model.onTransformUpdate (model.go:142-182)
    V3 position, scale;
    M4 rotation;
    decompose(M.transform, position, scale, rotation);
    V3 center = v3(0, 0, 0);
    if (!skinned)
      center = v3((mesh.dmax[0] + mesh.dmin[0]) / 2, (mesh.dmax[1] + mesh.dmin[1]) / 2, (mesh.dmax[2] + mesh.dmin[2]) / 2);
    center = mult_vec(rotation, center);
    position.x += center.x * scale.x;
    position.y += center.y * scale.y;
    position.z += center.z * scale.z;
    const V3 mn = v3(mesh.dmin[0] * scale.x, mesh.dmin[1] * scale.y, mesh.dmin[2] * scale.z);
    const V3 mx = v3(mesh.dmax[0] * scale.x, mesh.dmax[1] * scale.y, mesh.dmax[2] * scale.z);
    const float radius = magnitude(sub(mx, mn)) / 2;
    // SphereInFrustum (camera.go:933-983)
    const V3 cpos = v3(J.pos[0], J.pos[1], J.pos[2]);
    const V3 diff = sub(position, cpos);
    float pcZ = dot(diff, v3(-J.fwd[0], -J.fwd[1], -J.fwd[2]));
    const V3 cup = v3(J.up[0], J.up[1], J.up[2]);
    const V3 cright = v3(J.right[0], J.right[1], J.right[2]);
    if (pcZ > J.far_ + radius || pcZ < J.near_ - radius) {
      in = false;
    } else if (J.perspective) {
      float d = J.sf_y * radius;
      pcZ *= J.sf_tang;
      const float pcY = dot(diff, cup);
      if (pcY > pcZ + d || pcY < -pcZ - d) in = false;
      else {
        pcZ *= J.aspect;
        d = J.sf_x * radius;
        const float pcX = dot(diff, cright);
        if (pcX > pcZ + d || pcX < -pcZ - d) in = false;
      }
    } else {
      const float width = J.ortho_scale;
      const float height = width / J.aspect;
      const float pcY = dot(diff, cup);
      if (-height / 2 - radius > pcY || pcY > height / 2 + radius) in = false;
      else {
        const float pcX = dot(diff, cright);
        if (-width / 2 - radius > pcX || pcX > width / 2 + radius) in = false;
      }
    }
  }
  const float dist = dist_sq(v3(J.pos[0], J.pos[1], J.pos[2]), row3(M.transform));
  for (uint32_t p = 0; p < mesh.nparts; p++) {
    const uint32_t c = M.cand0 + p;
    const MaterialGPU* mat = stage_material(ms, M, p);
    const bool transparent = stage_transparent(M.color, mat);
    const bool drawn = in && !(mat && !(mat->bits & MB_VISIBLE));
    ms.cand[c] = (drawn ? 1u : 0u) | (transparent ? 2u : 0u) | (M.cam << 2);
    ms.cand_model[c] = m;
    ms.cand_dist[c] = dist;
    ms.cand_order[c] = mat ? mat->render_order : 0;
    const MeshPartGPU pr = mesh.parts[p];
    const PairSizes z = pair_sizes(ms, M, mesh, pr, mat, transparent);
    ms.cand_cnt[2 * c] = make_uint4(z.vcount, z.tcount, z.nvchunks, z.ntiles);
    ms.cand_cnt[2 * c + 1] = make_uint4(z.nsblocks, z.pscan, z.hist, z.lcount);
    ms.cand_bytes[c] = make_uint4(z.bytes_vertex, z.bytes_cull, z.bytes_uv, 0u);
    ms.cand_info[c] = make_uint4(M.index, p, pr.triangle_count, pr.triangle_start);
  }
}

// ---------------------------------------------------------------------------
// k_plist
// ---------------------------------------------------------------------------

constexpr uint32_t kListThreads = 1024;
constexpr uint32_t kListWaves = kListThreads / 64;

// Exclusive block-wide scan of N values per thread; every thread gets its
// exclusive prefixes and the block totals.  s_w holds kListWaves + 1 rows:
// each wave's total, scanned in place by the first wavefront, and the block
// total.
template <int N>
__device__ __forceinline__ void block_scan(uint32_t (&v)[N], uint32_t (&tot)[N], uint32_t (*s_w)[N]) {
  const uint32_t lane = threadIdx.x % 64u, wv = threadIdx.x / 64u;
  uint32_t inc[N];
#pragma unroll
  for (int k = 0; k < N; k++) inc[k] = v[k];
#pragma unroll
  for (int o = 1; o < 64; o <<= 1)
#pragma unroll
    for (int k = 0; k < N; k++) {
      const uint32_t y = __shfl_up(inc[k], o);
      if (lane >= (uint32_t)o) inc[k] += y;
    }
  if (lane == 63)
#pragma unroll
    for (int k = 0; k < N; k++) s_w[wv][k] = inc[k];
  __syncthreads();
  if (wv == 0) {
    uint32_t x[N], x0[N];
#pragma unroll
    for (int k = 0; k < N; k++) x[k] = x0[k] = lane < kListWaves ? s_w[lane][k] : 0u;
#pragma unroll
    for (int o = 1; o < (int)kListWaves; o <<= 1)
#pragma unroll
      for (int k = 0; k < N; k++) {
        const uint32_t y = __shfl_up(x[k], o);
        if (lane >= (uint32_t)o) x[k] += y;
      }
    if (lane < kListWaves)
#pragma unroll
      for (int k = 0; k < N; k++) s_w[lane][k] = x[k] - x0[k];
    if (lane == kListWaves - 1)
#pragma unroll
      for (int k = 0; k < N; k++) s_w[kListWaves][k] = x[k];
  }
  __syncthreads();
#pragma unroll
  for (int k = 0; k < N; k++) {
    v[k] = s_w[wv][k] + inc[k] - v[k];
    tot[k] = s_w[kListWaves][k];
  }
  __syncthreads();   // s_w is reused by the next scan
}

// Stable order of two transparent render pairs of one camera
// (camera.go:2050-2055): RenderOrder ascending, then distance² descending;
// candidates that compare equal keep candidate (traversal) order.
__device__ __forceinline__ bool t_before(int32_t oa, float da, uint32_t ia, int32_t ob, float db, uint32_t ib) {
  if (oa != ob) return oa < ob;
  if (da != db) return da > db;
  return ia < ib;
}

// In the first phase each thread takes kItems consecutive candidates per
// round (loaded coalesced through LDS), so a round of kListThreads * kItems
// candidates needs one block scan.
constexpr uint32_t kItems = 4;
constexpr uint32_t kRound = kListThreads * kItems;

__global__ __launch_bounds__(kListThreads, 1) void k_plist(FrameArgs a) {
  const ModelStage& ms = a.ms;
  const uint32_t tid = threadIdx.x;
  __shared__ uint32_t s_w2[kListWaves + 1][2];
  __shared__ uint32_t s_w8[kListWaves + 1][8];
  __shared__ uint32_t s_S[kListThreads], s_T[kListThreads];   // each thread's first prefix in the round
  __shared__ uint32_t s_stage[kRound];                          // a round's candidate flags
  __shared__ unsigned long long s_bytes[kListWaves][3];
  uint32_t* job_S0 = ms.job_s0;   // [njobs] solid and transparent drawn candidates before each job's first
  uint32_t* job_T0 = ms.job_t0;

  // Phase 1: solid render pairs in candidate order; drawn transparents listed (tlist)
  uint32_t cS = 0, cT = 0;
  for (uint32_t base = 0; base < ms.ncand; base += kRound) {
    const uint32_t c0 = base + tid * kItems;
#pragma unroll
    for (uint32_t i = 0; i < kItems; i++) {   // coalesced loads, then each thread takes kItems consecutive
      const uint32_t c = base + i * kListThreads + tid;
      s_stage[i * kListThreads + tid] = c < ms.ncand ? ms.cand[c] : 0u;
    }
    __syncthreads();
    uint32_t fl[kItems], job[kItems];
    uint32_t v[2] = {0, 0};
#pragma unroll
    for (uint32_t i = 0; i < kItems; i++) {
      const uint32_t w = s_stage[tid * kItems + i];
      fl[i] = w & 3u;
      job[i] = w >> 2;
      v[0] += fl[i] == 1u;
      v[1] += fl[i] == 3u;
    }
    uint32_t tot[2];
    block_scan<2>(v, tot, s_w2);
    s_S[tid] = cS + v[0];
    s_T[tid] = cT + v[1];
    __syncthreads();
    // prefixes at each job's first candidate inside this round
    for (uint32_t j = tid; j < ms.njobs; j += kListThreads) {
      const uint32_t jc = ms.jobs[j].cand0;
      if (jc >= base && jc < base + kRound && jc < ms.ncand) {
        const uint32_t t = (jc - base) / kItems;
        uint32_t S = s_S[t], T = s_T[t];
        for (uint32_t c = base + t * kItems; c < jc; c++) {
          const uint32_t f = s_stage[c - base] & 3u;
          S += f == 1u;
          T += f == 3u;
        }
        job_S0[j] = S;
        job_T0[j] = T;
      }
    }
    __syncthreads();
    uint32_t S = s_S[tid], T = s_T[tid];
#pragma unroll
    for (uint32_t i = 0; i < kItems; i++) {
      const uint32_t c = c0 + i;
      if (fl[i] == 1u) {
        ms.pair_cand[job_T0[job[i]] + S] = c;   // pairs of earlier jobs, then this job's solids before it
        S++;
      } else if (fl[i] == 3u) {
        ms.tlist[T] = c;
        T++;
      }
    }
    cS += tot[0];
    cT += tot[1];
    __syncthreads();   // s_S, s_T and s_stage are rewritten by the next round
  }
  for (uint32_t j = tid; j < ms.njobs; j += kListThreads)
    if (ms.jobs[j].cand0 >= ms.ncand) {   // no candidates from this job on
      job_S0[j] = cS;
      job_T0[j] = cT;
    }
  __syncthreads();
  const uint32_t nS = cS, nT = cT, P = nS + nT;

  // Phase 2: each transparent's place among its camera's transparents
  for (uint32_t i = tid; i < nT; i += kListThreads) {
    const uint32_t c = ms.tlist[i];
    const uint32_t j = ms.cand[c] >> 2;
    const uint32_t T0 = job_T0[j];
    const uint32_t T1 = j + 1 < ms.njobs ? job_T0[j + 1] : nT;
    const uint32_t S1 = j + 1 < ms.njobs ? job_S0[j + 1] : nS;
    const int32_t oi = ms.cand_order[c];
    const float di = ms.cand_dist[c];
    uint32_t rank = 0;
    for (uint32_t k = T0; k < T1; k++) {
      const uint32_t ck = ms.tlist[k];
      if (t_before(ms.cand_order[ck], ms.cand_dist[ck], k, oi, di, i)) rank++;
    }
    ms.pair_cand[S1 + T0 + rank] = c;   // earlier jobs' pairs and this job's solids, then its transparents
  }
  __syncthreads();

  // Phase 3: slices and prefix arrays over the pairs, one pair per thread
  // and round, so every load and store of a wavefront is contiguous
  uint32_t carry[8] = {0, 0, 0, 0, 0, 0, 0, 0};
  unsigned long long bv = 0, bc = 0, bu = 0;
  const size_t mp = ms.max_pairs;
  for (uint32_t base = 0; base < P; base += kListThreads) {
    const uint32_t p = base + tid;
    uint32_t v[8] = {0, 0, 0, 0, 0, 0, 0, 0};
    uint32_t c = 0;
    if (p < P) {
      c = ms.pair_cand[p];
      const uint4 q0 = ms.cand_cnt[2 * c], q1 = ms.cand_cnt[2 * c + 1];
      v[0] = q0.x; v[1] = q0.y; v[2] = q0.z; v[3] = q0.w;
      v[4] = q1.x; v[5] = q1.y; v[6] = q1.z; v[7] = q1.w;
      const uint4 b = ms.cand_bytes[c];
      bv += b.x;
      bc += b.y;
      bu += b.z;
    }
    uint32_t tot[8];
    block_scan<8>(v, tot, s_w8);
    if (p < P) {
      ms.pc[0 * mp + p] = carry[0] + v[0];   // first vertex record
      ms.pc[1 * mp + p] = carry[1] + v[1];   // first triangle slot
      ms.pc[6 * mp + p] = carry[6] + v[6];   // first histogram entry
      ms.pc[7 * mp + p] = carry[7] + v[7];   // first per-light value
      ms.vchunk_prefix[p] = carry[2] + v[2];
      ms.tile_prefix[p] = carry[3] + v[3];
      ms.sblock_prefix[p] = carry[4] + v[4];
      ms.pscan_prefix[p] = carry[5] + v[5];
      const uint4 info = ms.cand_info[c];
      PairOut o;
      o.model = info.x;
      o.part = info.y;
      o.triangle_count = info.z;
      o.tbase = carry[1] + v[1];
      o.out_base = 3ull * o.tbase;
      o.triangle_start = info.w;
      o.mrec = ms.cand_model[c];
      ms.pair_out[p] = o;
    }
#pragma unroll
    for (int k = 0; k < 8; k++) carry[k] += tot[k];
  }
  // past the last pair the prefix arrays hold the totals, so a search over
  // all max_pairs + 1 entries finds only actual pairs
  for (uint32_t p = P + tid; p <= ms.max_pairs; p += kListThreads) {
    ms.vchunk_prefix[p] = carry[2];
    ms.tile_prefix[p] = carry[3];
    ms.sblock_prefix[p] = carry[4];
    ms.pscan_prefix[p] = carry[5];
  }
  // algorithmic bytes of the frame (for the profile)
  for (int o = 32; o > 0; o >>= 1) {
    bv += __shfl_xor(bv, o);
    bc += __shfl_xor(bc, o);
    bu += __shfl_xor(bu, o);
  }
  if (tid % 64u == 0) {
    s_bytes[tid / 64u][0] = bv;
    s_bytes[tid / 64u][1] = bc;
    s_bytes[tid / 64u][2] = bu;
  }
  __syncthreads();
  if (tid == 0) {
    unsigned long long B[3] = {0, 0, 0};
    for (uint32_t w = 0; w < kListWaves; w++)
      for (int k = 0; k < 3; k++) B[k] += s_bytes[w][k];
    uint32_t* fc = a.counts;
    fc[FC_PAIRS] = P;
    fc[FC_VCHUNKS] = carry[2];
    fc[FC_TILES] = carry[3];
    fc[FC_SBLOCKS] = carry[4];
    fc[FC_PSCAN] = carry[5];
    fc[FC_TRANSPARENTS] = nT;
    fc[FC_SUMV] = carry[0];
    fc[FC_SUMT] = carry[1];
    fc[FC_BYTES_VERTEX_LO] = (uint32_t)B[0];
    fc[FC_BYTES_VERTEX_HI] = (uint32_t)(B[0] >> 32);
    fc[FC_BYTES_CULL_LO] = (uint32_t)B[1];
    fc[FC_BYTES_CULL_HI] = (uint32_t)(B[1] >> 32);
    fc[FC_BYTES_UV_LO] = (uint32_t)B[2];
    fc[FC_BYTES_UV_HI] = (uint32_t)(B[2] >> 32);
  }
  for (uint32_t j = tid; j < ms.njobs; j += kListThreads) ms.job_pairs[j] = job_S0[j] + job_T0[j];
  if (tid == 0) ms.job_pairs[ms.njobs] = P;
}

}  // namespace

hipError_t launch_model_stage(const FrameArgs& a, hipStream_t s) {
  if (a.ms.nmodels)
    hipLaunchKernelGGL(k_models, dim3((a.ms.nmodels + kModelThreads - 1) / kModelThreads), dim3(kModelThreads), 0, s, a);
  hipLaunchKernelGGL(k_plist, dim3(1), dim3(kListThreads), 0, s, a);
  return hipGetLastError();
}

}  // namespace t3d
