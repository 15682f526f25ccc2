// t3d_serve.cpp — the serving form: a render-request loop over resident
// meshes with HBM-budgeted admission and continuous batching.
//
// A job is one Camera.Render (camera.go:1269): a camera, a subset of the
// resident models and a light list.  A worker thread owns `slots` launch
// slots, each a HIP stream with a completion event.  Whenever a slot is free,
// every admitted job queued at that moment (FIFO, up to max_batch_jobs,
// within the budget) becomes one batch: one launch chain whose render pairs
// carry their job's camera index, so the batch's jobs share the kernels of
// the geometry path.  Batches on different slots overlap on the device; jobs
// arriving meanwhile wait for the next free slot.  A batch renders into a
// frame taken from a pool; when it completes the slot is free at once and
// the frame stays with the batch until every job of the batch is released,
// then returns to the pool.  A batched job's results are bit-identical to
// rendering it alone: the render pairs, uniforms and arithmetic of each job
// do not depend on the other jobs of its batch.
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
  std::vector<t3d_bone_pose> poses;
  std::vector<t3d_mat4> roots;     // armature parents of posed models
  std::vector<float> skw;
  std::vector<t3d_light_desc> group;
  std::vector<t3d_light_desc> nested;   // the volumes' use_lights
  t3d_render_desc desc{};
};

struct Batch;

struct Job {
  uint64_t id = 0;
  OwnedDesc d;
  std::vector<uint64_t> meshes;      // referenced mesh handles (refs held)
  std::vector<uint64_t> volumes;     // referenced light volumes (refs held)
  std::vector<uint64_t> animations;  // referenced animations (refs held)
  bool solo = false;                 // runs in a batch of its own (its batch failed to build with others)
  uint64_t est_bytes = 0;
  uint64_t triangles = 0;
  int32_t state = T3D_JOB_QUEUED;
  t3d_status status = T3D_OK;
  std::string error;
  double submit_ms = 0, start_ms = 0, done_ms = 0;
  Batch* batch = nullptr;            // while admitted and not released
  uint32_t index_in_batch = 0;
  uint64_t batch_id = 0;
  uint32_t batch_jobs = 0;
  uint64_t survivors = 0;
  uint32_t pairs = 0;
  bool released = false;
  int pins = 0;                      // downloads in progress
};

// One launch chain: its frame holds the outputs of all its jobs until the
// last of them is released.
struct Batch {
  uint64_t id = 0;
  t3d_frame* frame = nullptr;
  std::vector<uint64_t> jobs;
  uint32_t unreleased = 0;
  bool running = true;
};

struct Slot {
  hipStream_t stream = nullptr;
  hipEvent_t done = nullptr;
  Batch* batch = nullptr;            // the batch on the device, or null when free
  bool reserved = false;             // being launched outside the server lock
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
  std::unordered_map<uint64_t, std::unique_ptr<Batch>> batches_alive;
  std::vector<t3d_frame*> free_frames;   // frames of finished, fully released batches
  hipStream_t dl_stream = nullptr;       // downloads
  std::mutex dl_mu;
  // download packing buffer of t3d_job_download, shared by all jobs: grown under dl_mu and ctx->mu, its size
  // readable under mu (counted with the frames the server holds)
  t3d::DevBuf pack;
  uint64_t pack_cap = 0;
  std::atomic<uint64_t> pack_bytes{0};
  uint64_t wakeups = 0;                  // slot completions signalled from the slot streams
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
  // a volume's use_lights into o.nested (sized first: pointers are taken as it fills)
  auto nested_of = [](const t3d_light_desc& l) -> uint32_t {
    return (l.type == T3D_LIGHT_VOLUME && l.use_lights) ? l.use_light_count : 0u;
  };
  size_t nn = 0;
  for (const auto& l : o.lights) nn += nested_of(l);
  for (const auto& m : o.models)
    if (m.light_group_active && m.light_group)
      for (uint32_t k = 0; k < m.light_group_count; k++) nn += nested_of(m.light_group[k]);
  o.nested.reserve(nn);
  auto own_nested = [&](t3d_light_desc& l) {
    const uint32_t n = nested_of(l);
    if (!n) return;
    const size_t at = o.nested.size();
    o.nested.insert(o.nested.end(), l.use_lights, l.use_lights + n);
    for (size_t k = at; k < o.nested.size(); k++) {   // cells of cells are refused at build; keep no foreign pointer
      o.nested[k].use_lights = nullptr;
      o.nested[k].use_light_count = 0;
    }
    l.use_lights = o.nested.data() + at;
  };
  for (auto& l : o.lights) own_nested(l);
  // sizes first so the vectors never reallocate while pointers are taken
  size_t nm = 0, nb = 0, np = 0, nr = 0, nk = 0, ng = 0;
  for (const auto& m : o.models) {
    auto it = c->meshes.find(m.mesh);
    if (m.mesh && it == c->meshes.end()) return fail(c, T3D_ERR_NOT_FOUND, "submit: unknown mesh %llu", (unsigned long long)m.mesh);
    if (!m.mesh) continue;
    const Mesh& mesh = *it->second;
    if (m.materials) nm += mesh.parts.size();
    if (m.skinned && m.bone_influences) nb += m.bone_count;
    if (m.skinned && m.bone_poses) {
      np += m.bone_count;
      if (m.armature_parent) nr++;
    }
    if (m.shape_key_weights) nk += mesh.K;
    if (m.light_group_active && m.light_group) ng += m.light_group_count;
  }
  o.mats.reserve(nm);
  o.bones.reserve(nb);
  o.poses.reserve(np);
  o.roots.reserve(nr);
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
    if (m.skinned && m.bone_poses) {
      size_t at = o.poses.size();
      o.poses.insert(o.poses.end(), m.bone_poses, m.bone_poses + m.bone_count);
      m.bone_poses = o.poses.data() + at;
      if (m.armature_parent) {
        o.roots.push_back(*m.armature_parent);
        m.armature_parent = &o.roots.back();
      }
    } else {
      m.bone_poses = nullptr;
      m.armature_parent = nullptr;
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
      for (size_t k = at; k < o.group.size(); k++) own_nested(o.group[k]);
    }
    j.meshes.push_back(m.mesh);
    for (const auto& pr : mesh.parts) j.triangles += pr.triangle_count;
    if (m.skinned && m.bone_poses && m.animation) j.animations.push_back(m.animation);
    if (m.light_group_active && m.light_group)
      for (uint32_t k = 0; k < m.light_group_count; k++)
        if (m.light_group[k].type == T3D_LIGHT_VOLUME) j.volumes.push_back(m.light_group[k].volume);
  }
  for (const auto& l : o.lights)
    if (l.type == T3D_LIGHT_VOLUME) j.volumes.push_back(l.volume);
  o.desc.camera = &o.cam;
  o.desc.world = in->world ? &o.world : nullptr;
  o.desc.light_count = in->light_count;
  o.desc.lights = o.lights.data();
  o.desc.model_count = in->model_count;
  o.desc.models = o.models.data();
  o.desc.exec_mode = T3D_EXEC_BATCHED;
  o.desc.process_all_triangles = in->process_all_triangles;
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
    for (uint64_t h : j.volumes) {
      auto it = s->ctx->volumes.find(h);
      if (it != s->ctx->volumes.end()) it->second->refs--;
    }
    for (uint64_t h : j.animations) {
      auto it = s->ctx->animations.find(h);
      if (it != s->ctx->animations.end()) it->second->refs--;
    }
  }
  j.meshes.clear();
  j.volumes.clear();
  j.animations.clear();
}

// Called with ctx->mu held: takes the references drop_job gives back.
void hold_job(t3d_ctx* c, Job& j) {
  for (uint64_t h : j.meshes) c->meshes[h]->refs++;
  for (uint64_t h : j.volumes) c->volumes[h]->refs++;
  for (uint64_t h : j.animations) c->animations[h]->refs++;
}

uint64_t frames_held(const t3d_server* s) {
  uint64_t b = s->pack_bytes.load();
  for (const auto& kv : s->batches_alive) b += frame_capacity(kv.second->frame);
  for (const t3d_frame* f : s->free_frames) b += frame_capacity(f);
  return b;
}

// Called with s->mu held: a batch whose jobs are all released gives its frame
// back to the pool.
void retire_batch(t3d_server* s, Batch* b) {
  if (b->running || b->unreleased) return;
  s->free_frames.push_back(b->frame);
  const uint64_t id = b->id;
  s->batches_alive.erase(id);   // deletes b
}

// Called with s->mu held: frees a finished job's share of its batch.
void release_job(t3d_server* s, Job& j) {
  if (j.released) return;
  j.released = true;
  drop_job(s, j);
  s->inflight_bytes -= j.est_bytes;
  Batch* b = j.batch;
  j.batch = nullptr;
  if (b && b->unreleased > 0) {
    b->unreleased--;
    retire_batch(s, b);
  }
}

// Called with s->mu held, after the slot's event completed.
void complete_slot(t3d_server* s, int si) {
  Slot& sl = s->slots[si];
  Batch* b = sl.batch;
  sl.batch = nullptr;
  b->running = false;
  frame_collect(s->ctx, b->frame, false);
  const double now = s->now_ms();
  for (uint32_t k = 0; k < b->jobs.size(); k++) {
    Job& j = *s->jobs[b->jobs[k]];
    const auto& op = b->frame->out_pairs();
    const uint32_t p0 = b->frame->job_out_begin(k), p1 = b->frame->job_out_begin(k + 1);
    j.pairs = p1 - p0;
    j.survivors = 0;
    for (uint32_t p = p0; p < p1; p++) j.survivors += op[p].survivor_count;
    j.state = T3D_JOB_DONE;
    j.done_ms = now;
    s->completed++;
    s->tris_done += j.triangles;
    s->latencies.push_back(j.done_ms - j.submit_ms);
    s->last_done_ms = now;
  }
  if (s->cfg.auto_release)
    for (uint64_t id : std::vector<uint64_t>(b->jobs)) release_job(s, *s->jobs[id]);
  else
    retire_batch(s, b);
  s->cv_done.notify_all();
}

// Called with s->mu held: a batch of several jobs failed to build or
// launch.  Its jobs go back to the front of the queue, each to run in a
// batch of its own, so only a job that fails by itself fails.
void requeue_solo(t3d_server* s, Batch* b) {
  for (auto it = b->jobs.rbegin(); it != b->jobs.rend(); ++it) {
    Job& j = *s->jobs[*it];
    j.state = T3D_JOB_QUEUED;
    j.solo = true;
    j.batch = nullptr;
    j.batch_id = 0;
    j.batch_jobs = 0;
    s->inflight_bytes -= j.est_bytes;
    s->queue.push_front(j.id);
  }
  b->running = false;
  b->unreleased = 0;
  retire_batch(s, b);
}

// Slot stream callback after a batch's chain: wakes the worker, which then
// finds the slot's completion event done.  Runs on a runtime thread; takes
// the server lock and calls no HIP function.
void slot_done(void* arg) {
  t3d_server* s = static_cast<t3d_server*>(arg);
  {
    std::lock_guard<std::mutex> lk(s->mu);
    s->wakeups++;
  }
  s->cv_work.notify_one();
}

// Called with s->mu held: the batch never ran (or its launch failed).
void fail_batch(t3d_server* s, Batch* b, t3d_status st, const std::string& err) {
  for (uint64_t id : b->jobs) {
    Job& j = *s->jobs[id];
    j.state = T3D_JOB_FAILED;
    j.status = st;
    j.error = err;
    j.done_ms = s->now_ms();
    s->failed++;
    s->inflight_bytes -= j.est_bytes;
    j.batch = nullptr;
    j.released = true;
    drop_job(s, j);
  }
  b->running = false;
  b->unreleased = 0;
  retire_batch(s, b);
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
      if (!sl.batch || sl.reserved) continue;
      hipError_t q = hipEventQuery(sl.done);
      if (q == hipSuccess) {
        complete_slot(s, i);
        progressed = true;
      } else if (q == hipErrorNotReady) {
        running = true;
      } else {
        Batch* b = sl.batch;
        sl.batch = nullptr;
        fail_batch(s, b, T3D_ERR_DEVICE, hipGetErrorString(q));
        progressed = true;
      }
    }
    if (s->stop && !running && s->queue.empty()) break;
    // launch: every admitted job queued now goes into one batch on a free slot
    int free_slot = -1;
    for (int i = 0; i < (int)s->slots.size(); i++)
      if (!s->slots[i].batch && !s->slots[i].reserved) {
        free_slot = i;
        break;
      }
    if (free_slot >= 0 && !s->queue.empty() && !s->stop) {
      // A batch holds at most budget/slots of estimated frame memory (a
      // larger job still runs alone), so every slot's frame settles at a
      // fixed size instead of being reallocated as batch sizes vary.
      const uint64_t batch_cap = s->cfg.hbm_budget_bytes ? s->cfg.hbm_budget_bytes / s->slots.size() : UINT64_MAX;
      std::vector<uint64_t> take;
      uint64_t need = 0;
      while (!s->queue.empty() && take.size() < s->cfg.max_batch_jobs) {
        Job& j = *s->jobs[s->queue.front()];
        if (s->cfg.hbm_budget_bytes && s->inflight_bytes + j.est_bytes > s->cfg.hbm_budget_bytes) break;
        if (!take.empty() && (j.solo || need + j.est_bytes > batch_cap)) break;
        s->inflight_bytes += j.est_bytes;
        need += j.est_bytes;
        take.push_back(j.id);
        s->queue.pop_front();
        if (j.solo) break;
      }
      if (!take.empty()) {
        Slot& sl = s->slots[free_slot];
        auto bp = std::make_unique<Batch>();
        Batch* b = bp.get();
        b->id = s->next_batch++;
        b->jobs = take;
        b->unreleased = (uint32_t)take.size();
        // the frame with the most capacity from the pool, else a new one
        if (!s->free_frames.empty()) {
          auto it = std::max_element(s->free_frames.begin(), s->free_frames.end(),
                                     [](const t3d_frame* x, const t3d_frame* y) { return frame_capacity(x) < frame_capacity(y); });
          b->frame = *it;
          s->free_frames.erase(it);
        } else {
          b->frame = new t3d_frame();
        }
        // Pooled frames keep their buffers (sized with 25 % headroom).  When
        // the frames held plus this batch's need would pass that much over
        // the budget, idle pooled frames are freed.
        std::vector<t3d_frame*> trim;
        if (s->cfg.hbm_budget_bytes) {
          uint64_t held = frames_held(s);
          const uint64_t limit = s->cfg.hbm_budget_bytes + s->cfg.hbm_budget_bytes / 4;
          while (!s->free_frames.empty() && held + need + need / 4 > limit) {
            held -= frame_capacity(s->free_frames.back());
            trim.push_back(s->free_frames.back());
            s->free_frames.pop_back();
          }
        }
        std::vector<const t3d_render_desc*> descs;
        const double now = s->now_ms();
        for (uint32_t k = 0; k < take.size(); k++) {
          Job& j = *s->jobs[take[k]];
          j.state = T3D_JOB_RUNNING;
          j.start_ms = now;
          j.batch = b;
          j.index_in_batch = k;
          j.batch_id = b->id;
          j.batch_jobs = (uint32_t)take.size();
          descs.push_back(&j.d.desc);
        }
        s->batches_alive[b->id] = std::move(bp);
        sl.batch = b;
        sl.reserved = true;                // launched outside the server lock
        s->batches++;
        lk.unlock();
        t3d_status st;
        std::string err;
        {
          std::lock_guard<std::mutex> cl(s->ctx->mu);
          for (t3d_frame* f : trim) {
            frame_release(s->ctx, f);
            delete f;
          }
          st = frame_build(s->ctx, b->frame, descs.data(), (uint32_t)descs.size(), nullptr, true);
          if (st == T3D_OK) st = frame_launch(s->ctx, b->frame, sl.stream, sl.stream, T3D_EXEC_BATCHED, false, false);
          if (st != T3D_OK) {
            err = s->ctx->err;
            (void)hipStreamSynchronize(sl.stream);   // nothing of a failed launch stays in flight
            frame_release(s->ctx, b->frame);          // buffers sized for the failed batch go back at once
          }
        }
        if (st == T3D_OK && hipEventRecord(sl.done, sl.stream) != hipSuccess) {
          st = T3D_ERR_DEVICE;
          err = "hipEventRecord failed";
        }
        if (st == T3D_OK && hipLaunchHostFunc(sl.stream, slot_done, s) != hipSuccess) {
          st = T3D_ERR_DEVICE;
          err = "hipLaunchHostFunc failed";
          (void)hipStreamSynchronize(sl.stream);
        }
        lk.lock();
        sl.reserved = false;
        if (st != T3D_OK) {
          sl.batch = nullptr;
          if (take.size() > 1) requeue_solo(s, b);
          else fail_batch(s, b, st, err);
        }
        continue;
      }
    }
    if (!progressed) {
      // woken by submit, release, stop, or a slot stream's completion callback; the timeout is only a safety net
      if (running) s->cv_work.wait_for(lk, std::chrono::milliseconds(2));
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
  bool ok = hipStreamCreateWithFlags(&s->dl_stream, hipStreamNonBlocking) == hipSuccess;
  for (auto& sl : s->slots)
    ok = ok && hipStreamCreateWithFlags(&sl.stream, hipStreamNonBlocking) == hipSuccess &&
         hipEventCreateWithFlags(&sl.done, hipEventDisableTiming) == hipSuccess;
  if (!ok) {
    for (auto& x : s->slots) {
      if (x.stream) (void)hipStreamDestroy(x.stream);
      if (x.done) (void)hipEventDestroy(x.done);
    }
    if (s->dl_stream) (void)hipStreamDestroy(s->dl_stream);
    delete s;
    return fail(c, T3D_ERR_DEVICE, "server: stream/event creation failed");
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
  std::vector<t3d_frame*> frames;
  {
    std::lock_guard<std::mutex> lk(s->mu);
    for (auto& kv : s->jobs) release_job(s, *kv.second);
    for (auto& kv : s->batches_alive) frames.push_back(kv.second->frame);   // none once all are released
    s->batches_alive.clear();
    frames.insert(frames.end(), s->free_frames.begin(), s->free_frames.end());
    s->free_frames.clear();
  }
  {
    std::lock_guard<std::mutex> cl(s->ctx->mu);
    for (auto& sl : s->slots) {
      (void)hipStreamSynchronize(sl.stream);
      (void)hipEventDestroy(sl.done);
      (void)hipStreamDestroy(sl.stream);
    }
    (void)hipStreamSynchronize(s->dl_stream);
    (void)hipStreamDestroy(s->dl_stream);
    dev_free(s->ctx, s->pack);
    for (t3d_frame* f : frames) {
      frame_release(s->ctx, f);
      delete f;
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
    // every refusal a render of this descriptor alone would give, so a bad job never fails a batch
    st = validate_desc(s->ctx, &j->d.desc);
    if (st) return st;
    j->est_bytes = estimate_job_bytes(s->ctx, &j->d.desc);
    over_budget = s->cfg.hbm_budget_bytes && j->est_bytes > s->cfg.hbm_budget_bytes;
    if (over_budget) s->ctx->err = "submit: the job alone needs more frame memory than the server's HBM budget";
    else hold_job(s->ctx, *j);
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
  out->batch = j.batch_id;
  out->batch_jobs = j.batch_jobs;
  out->admitted_bytes = j.est_bytes;
  return T3D_OK;
}

}  // extern "C"

namespace {

// t3d_job_download and t3d_job_download_normals: the job's pair records and whichever arrays are asked for.
t3d_status job_download(t3d_server* s, uint64_t job, t3d_pair_result* pairs, t3d_vertex* cv, t3d_vertex* dv,
                        uint32_t* idx, uint32_t* tri_ids, t3d_vertex* nv) {
  if (!s) return T3D_ERR_INVALID_ARGUMENT;
  const t3d_frame* f = nullptr;
  uint32_t p0 = 0, p1 = 0;
  Job* jp = nullptr;
  {
    std::lock_guard<std::mutex> lk(s->mu);
    auto it = s->jobs.find(job);
    if (it == s->jobs.end()) return T3D_ERR_NOT_FOUND;
    Job& j = *it->second;
    if (j.state != T3D_JOB_DONE) return job_status(j) == T3D_OK ? T3D_ERR_NOT_READY : job_status(j);
    if (j.released || !j.batch) return sfail(s, T3D_ERR_NOT_FOUND, "job_download: the job's outputs were released");
    if (nv && !j.d.cam.render_normals)
      return sfail(s, T3D_ERR_NOT_FOUND, "job_download: the job's camera does not render normals");
    f = j.batch->frame;   // stays with the batch while this job is unreleased (pins block release)
    p0 = f->job_pair_begin[j.index_in_batch];
    p1 = f->job_pair_begin[j.index_in_batch + 1];
    if (pairs) {   // the job's flush records, offsets into its own concatenated arrays
      const auto& op = f->out_pairs();
      const uint32_t g0 = f->job_out_begin(j.index_in_batch), g1 = f->job_out_begin(j.index_in_batch + 1);
      uint64_t off = 0;
      for (uint32_t g = g0; g < g1; g++) {
        pairs[g - g0] = op[g];
        pairs[g - g0].vertex_offset = off;
        off += 3ull * pairs[g - g0].survivor_count;
      }
    }
    j.pins++;
    jp = &j;
  }
  t3d_status st = T3D_OK;
  if (cv || dv || idx || tri_ids || nv) {
    (void)hipSetDevice(s->ctx->device);
    std::lock_guard<std::mutex> dl(s->dl_mu);   // one download at a time on the download stream
    // the server's packing buffer: grown under ctx->mu (memory accounting), used under dl_mu only
    const uint64_t need = (cv || dv || idx || nv) ? download_pack_bytes(f, p0, p1, nv != nullptr) : 0;
    if (need > s->pack_cap) {
      std::lock_guard<std::mutex> cl(s->ctx->mu);
      st = ensure(s->ctx, s->pack, s->pack_cap, need);
      s->pack_bytes = s->pack.bytes;
    }
    if (!st) {
      // errors are reported into a local string (the worker may be writing ctx->err under ctx->mu)
      std::string err;
      st = frame_download_range(s->ctx, f, p0, p1, s->dl_stream, cv, dv, idx, tri_ids, nv, &s->pack, &s->pack_cap,
                                &err);
      if (st) {
        std::lock_guard<std::mutex> cl(s->ctx->mu);
        s->ctx->err = err;
      }
    }
  }
  std::lock_guard<std::mutex> lk(s->mu);
  jp->pins--;
  return st;
}

}  // namespace

extern "C" {

t3d_status t3d_job_download(t3d_server* s, uint64_t job, t3d_pair_result* pairs, t3d_vertex* cv, t3d_vertex* dv,
                            uint32_t* idx, uint32_t* tri_ids) {
  return job_download(s, job, pairs, cv, dv, idx, tri_ids, nullptr);
}

t3d_status t3d_job_download_normals(t3d_server* s, uint64_t job, t3d_vertex* nv) {
  if (!nv) return T3D_ERR_INVALID_ARGUMENT;
  return job_download(s, job, nullptr, nullptr, nullptr, nullptr, nullptr, nv);
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
    for (auto& sl : s->slots) busy |= sl.batch != nullptr;
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
  for (auto& sl : s->slots) out->running_batches += sl.batch != nullptr;
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
