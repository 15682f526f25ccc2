"""Serving form on the GPU: every job's results are the same alone and in a
continuously batched mix; admission by budget, cancellation and mesh
references behave as documented."""
import pytest

from tetra3d_amd import _abi as A
from tetra3d_amd import workloads as W
from tetra3d_amd.scene import Scene
from tests.parity import compare_frames

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def pool():
    models, jobs = W.serving_trace(pool_size=24, jobs=10, max_models=12, max_lights=6)
    return models, jobs, Scene("serving", W._world(0.2))


def job_models(models, job):
    return [models[i] for i in job.model_ids]


def test_jobs_same_alone_and_in_mix(gpu_ctx, oracle, pool):
    from tetra3d_amd.serving import Server
    models, jobs, scene = pool
    for m in models:
        gpu_ctx.upload_mesh(m.mesh)
    alone = []
    for j in jobs:
        alone.append(gpu_ctx.render(j.camera, scene, j.lights, job_models(models, j)))
    print("alone: survivors", [f.survivors for f in alone], flush=True)
    inputs = [gpu_ctx.inputs(j.camera, scene, j.lights, job_models(models, j)) for j in jobs]
    with Server(gpu_ctx, slots=2, max_batch_jobs=4) as srv:
        ids = [srv.submit_inputs(inp) for inp in inputs]
        for i in ids:
            srv.wait(i)
        batched = 0
        for k, i in enumerate(ids):
            r = srv.download(i)
            batched += r.info.batch_jobs > 1
            compare_frames(r, alone[k])
            srv.release(i)
        st = srv.stats()
        print("batches", st.batches, "jobs in shared batches", batched, "p50 %.3f p99 %.3f" % (st.p50_ms, st.p99_ms),
              flush=True)
        assert st.completed == len(jobs)
        assert batched > 0, "no job shared a batch"
    # and one of them against the oracle
    for m in models:
        if id(m.mesh) not in oracle._handles:
            oracle.register_mesh(m.mesh, m.mesh._handles["gpu"])
    ref = oracle.render(jobs[0].camera, scene, jobs[0].lights, job_models(models, jobs[0]))
    compare_frames(alone[0], ref)


def test_budget_limits_admission(gpu_ctx, pool):
    """With room for about two jobs, jobs still all complete, in more and
    smaller batches, and the bytes admitted at once never pass the budget."""
    from tetra3d_amd.serving import Server
    models, jobs, scene = pool
    for m in models:
        gpu_ctx.upload_mesh(m.mesh)
    inputs = [gpu_ctx.inputs(j.camera, scene, j.lights, job_models(models, j)) for j in jobs]
    with Server(gpu_ctx, slots=2, max_batch_jobs=8) as srv:
        ids = [srv.submit_inputs(inp) for inp in inputs]
        for i in ids:
            srv.wait(i)
        est = [srv.info(i).admitted_bytes for i in ids]
        free = [srv.download(i) for i in ids]
        for i in ids:
            srv.release(i)
    budget = 2 * max(est)
    with Server(gpu_ctx, slots=2, max_batch_jobs=8, hbm_budget=budget) as srv:
        ids = [srv.submit_inputs(inp) for inp in inputs]
        peak = 0
        for k, i in enumerate(ids):              # release as results are consumed
            srv.wait(i)
            peak = max(peak, srv.stats().memory_in_flight)
            compare_frames(srv.download(i), free[k])
            srv.release(i)
        st = srv.stats()
        print("budget", budget, "batches", st.batches, "peak admitted", peak, flush=True)
        assert st.completed == len(jobs) and peak <= budget


def test_budget_rejects_oversized_job(gpu_ctx, pool):
    from tetra3d_amd.runtime import T3DError
    from tetra3d_amd.serving import Server
    models, jobs, scene = pool
    for m in models:
        gpu_ctx.upload_mesh(m.mesh)
    with Server(gpu_ctx, slots=1, max_batch_jobs=2, hbm_budget=4096) as srv:
        with pytest.raises(T3DError) as e:
            srv.submit(jobs[0].camera, scene, jobs[0].lights, job_models(models, jobs[0]))
        assert e.value.status == A.T3D_ERR_BUDGET
        assert srv.stats().rejected == 1


def test_cancel_and_mesh_references(gpu_ctx, pool):
    from tetra3d_amd.runtime import T3DError
    from tetra3d_amd.serving import Server
    models, jobs, scene = pool
    for m in models:
        gpu_ctx.upload_mesh(m.mesh)
    with Server(gpu_ctx, slots=1, max_batch_jobs=1) as srv:
        inp = gpu_ctx.inputs(jobs[1].camera, scene, jobs[1].lights, job_models(models, jobs[1]))
        ids = [srv.submit_inputs(inp) for _ in range(12)]
        rc = srv.cancel(ids[-1])
        assert rc in (0, A.T3D_ERR_BUSY)
        # a mesh referenced by a submitted, unreleased job cannot be freed
        with pytest.raises(T3DError) as e:
            gpu_ctx.free_mesh(job_models(models, jobs[1])[0].mesh)
        assert e.value.status == A.T3D_ERR_BUSY
        for i in ids[:-1]:
            srv.wait(i)
        if rc == 0:
            assert srv.poll(ids[-1]) == A.T3D_ERR_CANCELLED
        for i in ids:
            srv.release(i)
        assert srv.stats().queue_depth == 0
