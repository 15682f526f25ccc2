// t3d_serve.cpp — the serving form: a render-request loop over resident
// meshes with HBM-budgeted admission and continuous batching.
//
// A job is one Camera.Render (camera.go:1269): a camera, a subset of the
// resident models and a light list.  A worker thread owns `slots` launch
// slots, each a frame (t3d_frame) with its own HIP stream.  Whenever a slot is
// free, every admitted job queued at that moment (FIFO, up to max_batch_jobs,
// within the budget) becomes one batch: one launch chain whose render pairs
// carry their job's camera index, so the batch's jobs share the kernels of
// the geometry path.  Batches on different slots overlap on the device; jobs
// arriving meanwhile wait for the next free slot.  A batched job's results
// are bit-identical to rendering it alone: the render pairs, uniforms and
// arithmetic of each job do not depend on the other jobs of its batch.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <chrono>
#include <cmath>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <thread>
#include <unordered_map>

#include "t3d_runtime.h"

using namespace t3d;

namespace {

using Clock = std::chrono::steady_clock;

// Deep copy of a t3d_render_desc; desc points into the owned vectors.
struct OwnedDesc {
  t3d_camera_desc cam{};
  t3d_world_desc world{};
  std::vector<t3d_light_desc> lights;
  std::vector<t3d_model_desc> models;
  std::vector<t3d_material_desc> mats;
  std::vector<t3d_mat4> bones;
  std::vector<float> skw;
  std::vector<t3d_light_desc> group;
  t3d_render_desc desc{};
};

struct Job {
  uint64_t id = 0;
  OwnedDesc d;
  std::vector<uint64_t> meshes;      // referenced mesh handles (refs held)
  uint64_t est_bytes = 0;
  uint64_t triangles = 0;
  int32_t state = T3D_JOB_QUEUED;
  t3d_status status = T3D_OK;
  std::string error;
  double submit_ms = 0, start_ms = 0, done_ms = 0;
  int slot = -1;
  uint32_t index_in_batch = 0;
  uint64_t batch = 0;
  uint32_t batch_jobs = 0;
  uint64_t survivors = 0;
  uint32_t pairs = 0;
  bool released = false;
  int pins = 0;                      // downloads in progress
};

// FREE: idle; RUNNING: batch on the device; HELD: finished, outputs kept until
// every job of the batch is released; TRIM: idle, buffers being given back.
enum SlotState { SLOT_FREE, SLOT_RUNNING, SLOT_HELD, SLOT_TRIM };

struct Slot {
  t3d_frame* frame = nullptr;
  hipStream_t stream = nullptr;
  hipEvent_t done = nullptr;
  SlotState state = SLOT_FREE;
  std::vector<uint64_t> jobs;
  uint32_t unreleased = 0;
  uint64_t batch = 0;
};

}  // namespace

struct t3d_server {
  t3d_ctx* ctx = nullptr;
  t3d_server_config cfg{};
  std::mutex mu;
  std::condition_variable cv_work;   // wakes the worker (submit, release, stop)
  std::condition_variable cv_done;   // wakes waiters
  std::deque<uint64_t> queue;
  std::unordered_map<uint64_t, std::unique_ptr<Job>> jobs;
  std::vector<Slot> slots;
  std::thread worker;
  bool stop = false;
  uint64_t next_id = 1, next_batch = 1;
  uint64_t inflight_bytes = 0;
  Clock::time_point t0;
  // statistics
  uint64_t submitted = 0, completed = 0, failed = 0, rejected = 0, cancelled = 0, batches = 0, tris_done = 0;
  std::vector<double> latencies;
  double first_submit_ms = -1, last_done_ms = 0;

  double now_ms() const { return std::chrono::duration<double, std::milli>(Clock::now() - t0).count(); }
};

namespace {

// Called with ctx->mu held: copies a descriptor and everything it points to.
t3d_status copy_desc(t3d_ctx* c, const t3d_render_desc* in, Job& j) {
  OwnedDesc& o = j.d;
  if (!in->camera) return fail(c, T3D_ERR_INVALID_ARGUMENT, "submit: no camera");
  if (in->light_count && !in->lights) return fail(c, T3D_ERR_INVALID_ARGUMENT, "submit: lights missing");
  if (in->model_count && !in->models) return fail(c, T3D_ERR_INVALID_ARGUMENT, "submit: models missing");
  o.cam = *in->camera;
  if (in->world) o.world = *in->world;
  o.lights.assign(in->lights, in->lights + in->light_count);
  o.models.assign(in->models, in->models + in->model_count);
  // sizes first so the vectors never reallocate while pointers are taken
  size_t nm = 0, nb = 0, nk = 0, ng = 0;
  for (const auto& m : o.models) {
    auto it = c->meshes.find(m.mesh);
    if (m.mesh && it == c->meshes.end()) return fail(c, T3D_ERR_NOT_FOUND, "submit: unknown mesh %llu", (unsigned long long)m.mesh);
    if (!m.mesh) continue;
    const Mesh& mesh = *it->second;
    if (m.materials) nm += mesh.parts.size();
    if (m.skinned && m.bone_influences) nb += m.bone_count;
    if (m.shape_key_weights) nk += mesh.K;
    if (m.light_group_active && m.light_group) ng += m.light_group_count;
  }
  o.mats.reserve(nm);
  o.bones.reserve(nb);
  o.skw.reserve(nk);
  o.group.reserve(ng);
  for (auto& m : o.models) {
    if (!m.mesh) continue;
    const Mesh& mesh = *c->meshes[m.mesh];
    if (m.materials) {
      size_t at = o.mats.size();
      o.mats.insert(o.mats.end(), m.materials, m.materials + mesh.parts.size());
      m.materials = o.mats.data() + at;
    }
    if (m.skinned && m.bone_influences) {
      size_t at = o.bones.size();
      o.bones.insert(o.bones.end(), m.bone_influences, m.bone_influences + m.bone_count);
      m.bone_influences = o.bones.data() + at;
    }
    if (m.shape_key_weights) {
      size_t at = o.skw.size();
      o.skw.insert(o.skw.end(), m.shape_key_weights, m.shape_key_weights + mesh.K);
      m.shape_key_weights = o.skw.data() + at;
    }
    if (m.light_group_active && m.light_group) {
      size_t at = o.group.size();
      o.group.insert(o.group.end(), m.light_group, m.light_group + m.light_group_count);
      m.light_group = o.group.data() + at;
    }
    j.meshes.push_back(m.mesh);
    for (const auto& pr : mesh.parts) j.triangles += pr.triangle_count;
  }
  o.desc.camera = &o.cam;
  o.desc.world = in->world ? &o.world : nullptr;
  o.desc.light_count = in->light_count;
  o.desc.lights = o.lights.data();
  o.desc.model_count = in->model_count;
  o.desc.models = o.models.data();
  o.desc.exec_mode = T3D_EXEC_BATCHED;
  return T3D_OK;
}

// Called with s->mu held.  Drops the job's mesh references (takes ctx->mu,
// lock order server -> context) and returns its admitted bytes.
void drop_job(t3d_server* s, Job& j) {
  {
    std::lock_guard<std::mutex> cl(s->ctx->mu);
    for (uint64_t h : j.meshes) {
      auto it = s->ctx->meshes.find(h);
      if (it != s->ctx->meshes.end()) it->second->refs--;
    }
  }
  j.meshes.clear();
}

// Called with s->mu held: frees a finished job's slot share.
void release_job(t3d_server* s, Job& j) {
  if (j.released) return;
  j.released = true;
  drop_job(s, j);
  if (j.slot >= 0) {
    Slot& sl = s->slots[j.slot];
    s->inflight_bytes -= j.est_bytes;
    if (sl.unreleased > 0 && --sl.unreleased == 0 && sl.state == SLOT_HELD) {
      sl.state = SLOT_FREE;
      sl.jobs.clear();
    }
  }
}

// Called with s->mu held, after the slot's event completed.
void complete_slot(t3d_server* s, int si) {
  Slot& sl = s->slots[si];
  frame_collect(s->ctx, sl.frame, false);
  const double now = s->now_ms();
  for (uint32_t k = 0; k < sl.jobs.size(); k++) {
    Job& j = *s->jobs[sl.jobs[k]];
    const uint32_t p0 = sl.frame->job_pair_begin[k], p1 = sl.frame->job_pair_begin[k + 1];
    j.pairs = p1 - p0;
    j.survivors = 0;
    for (uint32_t p = p0; p < p1; p++) j.survivors += sl.frame->pairs[p].survivor_count;
    j.state = T3D_JOB_DONE;
    j.done_ms = now;
    s->completed++;
    s->tris_done += j.triangles;
    s->latencies.push_back(j.done_ms - j.submit_ms);
    s->last_done_ms = now;
  }
  sl.state = SLOT_HELD;
  if (s->cfg.auto_release)
    for (uint64_t id : std::vector<uint64_t>(sl.jobs)) release_job(s, *s->jobs[id]);
  s->cv_done.notify_all();
}

void fail_batch(t3d_server* s, int si, t3d_status st, const std::string& err) {
  Slot& sl = s->slots[si];
  for (uint64_t id : sl.jobs) {
    Job& j = *s->jobs[id];
    j.state = T3D_JOB_FAILED;
    j.status = st;
    j.error = err;
    j.done_ms = s->now_ms();
    s->failed++;
    s->inflight_bytes -= j.est_bytes;
    j.slot = -1;
    j.released = true;
    drop_job(s, j);
  }
  sl.jobs.clear();
  sl.unreleased = 0;
  sl.state = SLOT_FREE;
  s->cv_done.notify_all();
}

void worker_main(t3d_server* s) {
  (void)hipSetDevice(s->ctx->device);
  std::unique_lock<std::mutex> lk(s->mu);
  for (;;) {
    bool progressed = false;
    bool running = false;
    for (int i = 0; i < (int)s->slots.size(); i++) {
      Slot& sl = s->slots[i];
      if (sl.state != SLOT_RUNNING) continue;
      hipError_t q = hipEventQuery(sl.done);
      if (q == hipSuccess) {
        complete_slot(s, i);
        progressed = true;
      } else if (q == hipErrorNotReady) {
        running = true;
      } else {
        fail_batch(s, i, T3D_ERR_DEVICE, hipGetErrorString(q));
        progressed = true;
      }
    }
    if (s->stop && !running && s->queue.empty()) break;
    // launch: every admitted job queued now goes into one batch on a free slot
    int free_slot = -1;
    for (int i = 0; i < (int)s->slots.size(); i++)
      if (s->slots[i].state == SLOT_FREE) {
        free_slot = i;
        break;
      }
    if (free_slot >= 0 && !s->queue.empty() && !s->stop) {
      Slot& sl = s->slots[free_slot];
      sl.jobs.clear();
      while (!s->queue.empty() && sl.jobs.size() < s->cfg.max_batch_jobs) {
        Job& j = *s->jobs[s->queue.front()];
        if (s->cfg.hbm_budget_bytes && s->inflight_bytes + j.est_bytes > s->cfg.hbm_budget_bytes) break;
        s->inflight_bytes += j.est_bytes;
        sl.jobs.push_back(j.id);
        s->queue.pop_front();
      }
      if (!sl.jobs.empty()) {
        const uint64_t batch = s->next_batch++;
        std::vector<const t3d_render_desc*> descs;
        const double now = s->now_ms();
        for (uint32_t k = 0; k < sl.jobs.size(); k++) {
          Job& j = *s->jobs[sl.jobs[k]];
          j.state = T3D_JOB_RUNNING;
          j.start_ms = now;
          j.slot = free_slot;
          j.index_in_batch = k;
          j.batch = batch;
          j.batch_jobs = (uint32_t)sl.jobs.size();
          descs.push_back(&j.d.desc);
        }
        sl.state = SLOT_RUNNING;   // reserved while building outside the server lock
        sl.unreleased = (uint32_t)sl.jobs.size();
        sl.batch = batch;
        s->batches++;
        // Frames keep their buffers between batches.  When the frames held by
        // the other slots plus this batch's need would pass the budget, the
        // buffers of idle slots are given back first.
        std::vector<int> trim;
        if (s->cfg.hbm_budget_bytes) {
          uint64_t need = 0, held = 0;
          for (uint64_t id : sl.jobs) need += s->jobs[id]->est_bytes;
          for (int i = 0; i < (int)s->slots.size(); i++)
            if (i != free_slot) held += frame_capacity(s->slots[i].frame);
          for (int i = 0; i < (int)s->slots.size() && held + need + need / 4 > s->cfg.hbm_budget_bytes; i++)
            if (i != free_slot && s->slots[i].state == SLOT_FREE && frame_capacity(s->slots[i].frame)) {
              held -= frame_capacity(s->slots[i].frame);
              trim.push_back(i);
              s->slots[i].state = SLOT_TRIM;   // kept out of launches while trimmed outside the lock
            }
        }
        lk.unlock();
        t3d_status st;
        std::string err;
        {
          std::lock_guard<std::mutex> cl(s->ctx->mu);
          for (int i : trim) frame_release(s->ctx, s->slots[i].frame);
          st = frame_build(s->ctx, sl.frame, descs.data(), (uint32_t)descs.size());
          if (st == T3D_OK) st = frame_launch(s->ctx, sl.frame, sl.stream, sl.stream, T3D_EXEC_BATCHED, false, false);
          if (st != T3D_OK) err = s->ctx->err;
        }
        if (st == T3D_OK && hipEventRecord(sl.done, sl.stream) != hipSuccess) {
          st = T3D_ERR_DEVICE;
          err = "hipEventRecord failed";
        }
        lk.lock();
        for (int i : trim) s->slots[i].state = SLOT_FREE;
        if (st != T3D_OK) fail_batch(s, free_slot, st, err);
        continue;
      }
    }
    if (!progressed) {
      if (running) s->cv_work.wait_for(lk, std::chrono::microseconds(20));
      else s->cv_work.wait(lk);
    }
  }
}

t3d_status sfail(t3d_server* s, t3d_status st, const char* msg) {
  if (s && s->ctx) {
    std::lock_guard<std::mutex> cl(s->ctx->mu);
    s->ctx->err = msg;
  }
  return st;
}

}  // namespace

extern "C" {

t3d_status t3d_server_create(t3d_ctx* c, const t3d_server_config* cfg, t3d_server** out) {
  if (!c || !cfg || !out) return T3D_ERR_INVALID_ARGUMENT;
  if (cfg->slots == 0 || cfg->slots > 16 || cfg->max_batch_jobs == 0)
    return fail(c, T3D_ERR_INVALID_ARGUMENT, "server: slots in [1, 16], max_batch_jobs >= 1");
  (void)hipSetDevice(c->device);
  auto* s = new t3d_server();
  s->ctx = c;
  s->cfg = *cfg;
  s->t0 = Clock::now();
  s->slots.resize(cfg->slots);
  for (auto& sl : s->slots) {
    sl.frame = new t3d_frame();
    if (hipStreamCreateWithFlags(&sl.stream, hipStreamNonBlocking) != hipSuccess ||
        hipEventCreateWithFlags(&sl.done, hipEventDisableTiming) != hipSuccess) {
      for (auto& x : s->slots) {
        if (x.stream) (void)hipStreamDestroy(x.stream);
        if (x.done) (void)hipEventDestroy(x.done);
        delete x.frame;
      }
      delete s;
      return fail(c, T3D_ERR_DEVICE, "server: stream/event creation failed");
    }
  }
  s->worker = std::thread(worker_main, s);
  *out = s;
  return T3D_OK;
}

t3d_status t3d_server_destroy(t3d_server* s) {
  if (!s) return T3D_ERR_INVALID_ARGUMENT;
  {
    std::lock_guard<std::mutex> lk(s->mu);
    for (uint64_t id : s->queue) {
      Job& j = *s->jobs[id];
      j.state = T3D_JOB_CANCELLED;
      j.released = true;
      drop_job(s, j);
    }
    s->queue.clear();
    s->stop = true;
  }
  s->cv_work.notify_all();
  s->worker.join();
  {
    std::lock_guard<std::mutex> lk(s->mu);
    for (auto& kv : s->jobs) release_job(s, *kv.second);
  }
  {
    std::lock_guard<std::mutex> cl(s->ctx->mu);
    for (auto& sl : s->slots) {
      (void)hipStreamSynchronize(sl.stream);
      frame_release(s->ctx, sl.frame);
      delete sl.frame;
      (void)hipEventDestroy(sl.done);
      (void)hipStreamDestroy(sl.stream);
    }
  }
  delete s;
  return T3D_OK;
}

t3d_status t3d_job_submit(t3d_server* s, const t3d_render_desc* desc, uint64_t* out_job) {
  if (!s || !desc || !out_job) return T3D_ERR_INVALID_ARGUMENT;
  *out_job = 0;
  auto j = std::make_unique<Job>();
  bool over_budget = false;
  {
    std::lock_guard<std::mutex> cl(s->ctx->mu);
    t3d_status st = copy_desc(s->ctx, desc, *j);
    if (st) return st;
    j->est_bytes = estimate_job_bytes(s->ctx, &j->d.desc);
    over_budget = s->cfg.hbm_budget_bytes && j->est_bytes > s->cfg.hbm_budget_bytes;
    if (over_budget) s->ctx->err = "submit: the job alone needs more frame memory than the server's HBM budget";
    else
      for (uint64_t h : j->meshes) s->ctx->meshes[h]->refs++;
  }
  // the server lock is taken only after the context lock is dropped (the
  // worker nests them server -> context)
  std::lock_guard<std::mutex> lk(s->mu);
  if (over_budget) {
    s->rejected++;
    return T3D_ERR_BUDGET;
  }
  if (s->stop) {
    drop_job(s, *j);
    return T3D_ERR_CANCELLED;
  }
  j->id = s->next_id++;
  j->submit_ms = s->now_ms();
  if (s->first_submit_ms < 0) s->first_submit_ms = j->submit_ms;
  s->submitted++;
  *out_job = j->id;
  s->queue.push_back(j->id);
  s->jobs[j->id] = std::move(j);
  s->cv_work.notify_one();
  return T3D_OK;
}

static t3d_status job_status(const Job& j) {
  switch (j.state) {
    case T3D_JOB_DONE: return T3D_OK;
    case T3D_JOB_FAILED: return j.status;
    case T3D_JOB_CANCELLED: return T3D_ERR_CANCELLED;
    default: return T3D_ERR_NOT_READY;
  }
}

t3d_status t3d_job_poll(t3d_server* s, uint64_t job) {
  if (!s) return T3D_ERR_INVALID_ARGUMENT;
  std::lock_guard<std::mutex> lk(s->mu);
  auto it = s->jobs.find(job);
  if (it == s->jobs.end()) return T3D_ERR_NOT_FOUND;
  return job_status(*it->second);
}

t3d_status t3d_job_wait(t3d_server* s, uint64_t job, uint32_t timeout_ms) {
  if (!s) return T3D_ERR_INVALID_ARGUMENT;
  std::unique_lock<std::mutex> lk(s->mu);
  auto deadline = Clock::now() + std::chrono::milliseconds(timeout_ms);
  for (;;) {
    auto it = s->jobs.find(job);
    if (it == s->jobs.end()) return T3D_ERR_NOT_FOUND;
    t3d_status st = job_status(*it->second);
    if (st != T3D_ERR_NOT_READY) return st;
    if (s->cv_done.wait_until(lk, deadline) == std::cv_status::timeout) {
      it = s->jobs.find(job);
      return it == s->jobs.end() ? T3D_ERR_NOT_FOUND : job_status(*it->second);
    }
  }
}

t3d_status t3d_job_info_get(t3d_server* s, uint64_t job, t3d_job_info* out) {
  if (!s || !out) return T3D_ERR_INVALID_ARGUMENT;
  std::lock_guard<std::mutex> lk(s->mu);
  auto it = s->jobs.find(job);
  if (it == s->jobs.end()) return T3D_ERR_NOT_FOUND;
  const Job& j = *it->second;
  std::memset(out, 0, sizeof *out);
  out->state = j.state;
  out->status = j.status;
  out->submit_ms = j.submit_ms;
  out->start_ms = j.start_ms;
  out->done_ms = j.done_ms;
  out->pair_count = j.pairs;
  out->triangles = j.triangles;
  out->survivors = j.survivors;
  out->batch = j.batch;
  out->batch_jobs = j.batch_jobs;
  return T3D_OK;
}

t3d_status t3d_job_download(t3d_server* s, uint64_t job, t3d_pair_result* pairs, t3d_vertex* cv, t3d_vertex* dv,
                            uint32_t* idx, uint32_t* tri_ids) {
  if (!s) return T3D_ERR_INVALID_ARGUMENT;
  Slot* sl = nullptr;
  uint32_t p0 = 0, p1 = 0;
  Job* jp = nullptr;
  {
    std::lock_guard<std::mutex> lk(s->mu);
    auto it = s->jobs.find(job);
    if (it == s->jobs.end()) return T3D_ERR_NOT_FOUND;
    Job& j = *it->second;
    if (j.state != T3D_JOB_DONE) return job_status(j) == T3D_OK ? T3D_ERR_NOT_READY : job_status(j);
    if (j.released) return sfail(s, T3D_ERR_NOT_FOUND, "job_download: the job's outputs were released");
    sl = &s->slots[j.slot];
    p0 = sl->frame->job_pair_begin[j.index_in_batch];
    p1 = sl->frame->job_pair_begin[j.index_in_batch + 1];
    if (pairs) {
      uint64_t off = 0;
      for (uint32_t p = p0; p < p1; p++) {
        pairs[p - p0] = sl->frame->pairs[p];
        pairs[p - p0].vertex_offset = off;
        off += 3ull * pairs[p - p0].survivor_count;
      }
    }
    j.pins++;
    jp = &j;
  }
  t3d_status st = T3D_OK;
  if (cv || dv || idx || tri_ids) {
    (void)hipSetDevice(s->ctx->device);
    st = frame_download_range(s->ctx, sl->frame, p0, p1, sl->stream, cv, dv, idx, tri_ids);
  }
  std::lock_guard<std::mutex> lk(s->mu);
  jp->pins--;
  return st;
}

t3d_status t3d_job_release(t3d_server* s, uint64_t job) {
  if (!s) return T3D_ERR_INVALID_ARGUMENT;
  std::lock_guard<std::mutex> lk(s->mu);
  auto it = s->jobs.find(job);
  if (it == s->jobs.end()) return T3D_ERR_NOT_FOUND;
  Job& j = *it->second;
  if (j.state == T3D_JOB_QUEUED || j.state == T3D_JOB_RUNNING) return T3D_ERR_BUSY;
  if (j.pins) return T3D_ERR_BUSY;
  release_job(s, j);
  s->jobs.erase(it);
  s->cv_work.notify_one();
  return T3D_OK;
}

t3d_status t3d_job_cancel(t3d_server* s, uint64_t job) {
  if (!s) return T3D_ERR_INVALID_ARGUMENT;
  std::lock_guard<std::mutex> lk(s->mu);
  auto it = s->jobs.find(job);
  if (it == s->jobs.end()) return T3D_ERR_NOT_FOUND;
  Job& j = *it->second;
  if (j.state == T3D_JOB_CANCELLED) return T3D_OK;
  if (j.state != T3D_JOB_QUEUED) return T3D_ERR_BUSY;   // running or finished
  s->queue.erase(std::find(s->queue.begin(), s->queue.end(), job));
  j.state = T3D_JOB_CANCELLED;
  j.released = true;
  drop_job(s, j);
  s->cancelled++;
  s->cv_done.notify_all();
  return T3D_OK;
}

t3d_status t3d_server_drain(t3d_server* s, uint32_t timeout_ms) {
  if (!s) return T3D_ERR_INVALID_ARGUMENT;
  std::unique_lock<std::mutex> lk(s->mu);
  auto deadline = Clock::now() + std::chrono::milliseconds(timeout_ms);
  for (;;) {
    bool busy = !s->queue.empty();
    for (auto& sl : s->slots) busy |= sl.state == SLOT_RUNNING;
    if (!busy) return T3D_OK;
    if (s->cv_done.wait_until(lk, deadline) == std::cv_status::timeout) return T3D_ERR_NOT_READY;
  }
}

t3d_status t3d_server_stats_get(t3d_server* s, t3d_server_stats* out) {
  if (!s || !out) return T3D_ERR_INVALID_ARGUMENT;
  std::lock_guard<std::mutex> lk(s->mu);
  std::memset(out, 0, sizeof *out);
  out->submitted = s->submitted;
  out->completed = s->completed;
  out->failed = s->failed;
  out->rejected = s->rejected;
  out->cancelled = s->cancelled;
  out->batches = s->batches;
  out->triangles = s->tris_done;
  out->queue_depth = (uint32_t)s->queue.size();
  for (auto& sl : s->slots) out->running_batches += sl.state == SLOT_RUNNING;
  out->memory_in_flight = s->inflight_bytes;
  if (!s->latencies.empty()) {
    std::vector<double> v = s->latencies;
    auto pct = [&](double q) {
      size_t k = std::min(v.size() - 1, (size_t)std::ceil(q * v.size()) - (q > 0 ? 1 : 0));
      std::nth_element(v.begin(), v.begin() + k, v.end());
      return v[k];
    };
    out->p50_ms = pct(0.50);
    out->p99_ms = pct(0.99);
  }
  out->first_submit_ms = s->first_submit_ms;
  out->last_done_ms = s->last_done_ms;
  return T3D_OK;
}

t3d_status t3d_server_stats_reset(t3d_server* s) {
  if (!s) return T3D_ERR_INVALID_ARGUMENT;
  std::lock_guard<std::mutex> lk(s->mu);
  s->submitted = s->completed = s->failed = s->rejected = s->cancelled = s->batches = s->tris_done = 0;
  s->latencies.clear();
  s->first_submit_ms = -1;
  s->last_done_ms = 0;
  return T3D_OK;
}

}  // extern "C"
