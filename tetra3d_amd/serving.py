"""ctypes binding of the serving form (t3d_server_* / t3d_job_* in include/t3d.h).

A Server keeps its context's meshes resident, admits jobs (one Camera.Render
each: camera + model subset + lights) by an HBM budget and batches them
continuously on several HIP streams.
"""
import ctypes as C

import numpy as np

from . import _abi as A
from .runtime import T3DError, lib as _lib

JOB_QUEUED, JOB_RUNNING, JOB_DONE, JOB_FAILED, JOB_CANCELLED = 0, 1, 2, 3, 4


class ServerConfig(C.Structure):
    _fields_ = [
        ("slots", C.c_uint32),
        ("max_batch_jobs", C.c_uint32),
        ("hbm_budget_bytes", C.c_uint64),
        ("auto_release", C.c_int32),
    ]


class JobInfo(C.Structure):
    _fields_ = [
        ("state", C.c_int32),
        ("status", C.c_int32),
        ("submit_ms", C.c_double),
        ("start_ms", C.c_double),
        ("done_ms", C.c_double),
        ("pair_count", C.c_uint32),
        ("triangles", C.c_uint64),
        ("survivors", C.c_uint64),
        ("batch", C.c_uint64),
        ("batch_jobs", C.c_uint32),
        ("admitted_bytes", C.c_uint64),
    ]


class ServerStats(C.Structure):
    _fields_ = [
        ("submitted", C.c_uint64),
        ("completed", C.c_uint64),
        ("failed", C.c_uint64),
        ("rejected", C.c_uint64),
        ("cancelled", C.c_uint64),
        ("batches", C.c_uint64),
        ("triangles", C.c_uint64),
        ("queue_depth", C.c_uint32),
        ("running_batches", C.c_uint32),
        ("memory_in_flight", C.c_uint64),
        ("p50_ms", C.c_double),
        ("p99_ms", C.c_double),
        ("first_submit_ms", C.c_double),
        ("last_done_ms", C.c_double),
    ]


_bound = False


def _bind():
    global _bound
    if _bound:
        return
    L = _lib()
    P, vp = C.POINTER, C.c_void_p
    L.t3d_server_create.argtypes = [vp, P(ServerConfig), P(vp)]
    L.t3d_server_destroy.argtypes = [vp]
    L.t3d_job_submit.argtypes = [vp, P(A.RenderDesc), P(C.c_uint64)]
    L.t3d_job_poll.argtypes = [vp, C.c_uint64]
    L.t3d_job_wait.argtypes = [vp, C.c_uint64, C.c_uint32]
    L.t3d_job_info_get.argtypes = [vp, C.c_uint64, P(JobInfo)]
    L.t3d_job_download.argtypes = [vp, C.c_uint64, P(A.PairResult), P(A.Vertex), P(A.Vertex), P(C.c_uint32),
                                   P(C.c_uint32)]
    L.t3d_job_release.argtypes = [vp, C.c_uint64]
    L.t3d_job_download_normals.argtypes = [vp, C.c_uint64, P(A.Vertex)]
    L.t3d_job_cancel.argtypes = [vp, C.c_uint64]
    L.t3d_server_drain.argtypes = [vp, C.c_uint32]
    L.t3d_server_stats_get.argtypes = [vp, P(ServerStats)]
    L.t3d_server_stats_reset.argtypes = [vp]
    for n in ("t3d_server_create", "t3d_server_destroy", "t3d_job_submit", "t3d_job_poll", "t3d_job_wait",
              "t3d_job_info_get", "t3d_job_download", "t3d_job_release", "t3d_job_cancel", "t3d_server_drain",
              "t3d_server_stats_get", "t3d_server_stats_reset", "t3d_job_download_normals"):
        getattr(L, n).restype = C.c_int32
    _bound = True


class JobResult:
    def __init__(self, pairs, color, depth, indices, triangle_ids, info, normals=None):
        self.normals = normals
        self.pairs = pairs
        self.color = color
        self.depth = depth
        self.indices = indices
        self.triangle_ids = triangle_ids
        self.info = info

    @property
    def survivors(self):
        return int(sum(p[3] for p in self.pairs))


class Server:
    """A render-request loop on a runtime.Context (whose meshes it renders)."""

    def __init__(self, ctx, slots=4, max_batch_jobs=16, hbm_budget=0, auto_release=False):
        _bind()
        self.L = _lib()
        self.ctx = ctx
        cfg = ServerConfig(slots, max_batch_jobs, hbm_budget, int(auto_release))
        self.srv = C.c_void_p()
        ctx._check(self.L.t3d_server_create(ctx.ctx, C.byref(cfg), C.byref(self.srv)))

    def _check(self, rc):
        return self.ctx._check(rc)

    def close(self):
        if self.srv:
            self.L.t3d_server_destroy(self.srv)
            self.srv = None

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def submit_inputs(self, inp):
        """Submit a prepared runtime RenderInputs; returns the job id."""
        job = C.c_uint64()
        self._check(self.L.t3d_job_submit(self.srv, C.byref(inp.render), C.byref(job)))
        return job.value

    def submit(self, camera, scene, lights, models):
        return self.submit_inputs(self.ctx.inputs(camera, scene, lights, models))

    def poll(self, job):
        return self.L.t3d_job_poll(self.srv, job)

    def wait(self, job, timeout_ms=60000):
        rc = self.L.t3d_job_wait(self.srv, job, timeout_ms)
        if rc == A.T3D_ERR_NOT_READY:
            raise TimeoutError("job %d not done after %d ms" % (job, timeout_ms))
        return self._check(rc)

    def info(self, job):
        i = JobInfo()
        self._check(self.L.t3d_job_info_get(self.srv, job, C.byref(i)))
        return i

    def download(self, job, triangles=True, normals=False):
        i = self.info(job)
        pairs = (A.PairResult * max(1, i.pair_count))()
        S = int(i.survivors)
        color = np.zeros((3 * S, 12), dtype=np.float32)
        depth = np.zeros((3 * S, 12), dtype=np.float32)
        idx = np.zeros(3 * S, dtype=np.uint32)
        tri = np.zeros(max(S, 1), dtype=np.uint32)
        VP = C.POINTER(A.Vertex)
        U32 = C.POINTER(C.c_uint32)
        self._check(self.L.t3d_job_download(self.srv, job, pairs, color.ctypes.data_as(VP), depth.ctypes.data_as(VP),
                                            idx.ctypes.data_as(U32), tri.ctypes.data_as(U32) if triangles else None))
        pl = [(p.model_index, p.part_index, p.triangle_count, p.survivor_count, p.vertex_offset)
              for p in pairs[:i.pair_count]]
        nrm = None
        if normals:
            nrm = np.zeros((3 * S, 12), dtype=np.float32)
            self._check(self.L.t3d_job_download_normals(self.srv, job, nrm.ctypes.data_as(VP)))
        return JobResult(pl, color, depth, idx, tri[:S] if triangles else None, i, nrm)

    def release(self, job):
        self._check(self.L.t3d_job_release(self.srv, job))

    def cancel(self, job):
        return self.L.t3d_job_cancel(self.srv, job)

    def drain(self, timeout_ms=120000):
        rc = self.L.t3d_server_drain(self.srv, timeout_ms)
        if rc != 0:
            raise T3DError(rc, "drain timed out")

    def stats(self):
        s = ServerStats()
        self._check(self.L.t3d_server_stats_get(self.srv, C.byref(s)))
        return s

    def reset_stats(self):
        self._check(self.L.t3d_server_stats_reset(self.srv))
