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


def test_invalid_job_refused_at_submit_mix_unaffected(gpu_ctx, oracle, pool):
    """A job that rendering alone would refuse is refused at submit with the
    same status, so it never joins a batch; the valid jobs submitted around
    it batch together and match their alone renders bit for bit."""
    from tetra3d_amd.runtime import T3DError
    from tetra3d_amd.serving import Server
    models, jobs, scene = pool
    for m in models:
        gpu_ctx.upload_mesh(m.mesh)
    good = [gpu_ctx.inputs(j.camera, scene, j.lights, job_models(models, j)) for j in jobs[:6]]
    alone = [gpu_ctx.render(j.camera, scene, j.lights, job_models(models, j)) for j in jobs[:6]]

    def custom_depth(inp):
        mat = inp.render.models[0].materials[0]
        mat.present = 1
        mat.visible = 1
        mat.custom_depth_function = 1

    def bad_light(inp):
        inp.render.lights[0].type = 99

    def go_callback(inp):
        inp.render.models[0].has_vertex_transform_function = 1

    cases = [(custom_depth, A.T3D_ERR_UNSUPPORTED), (bad_light, A.T3D_ERR_INVALID_ARGUMENT),
             (go_callback, A.T3D_ERR_UNSUPPORTED)]
    with Server(gpu_ctx, slots=1, max_batch_jobs=16) as srv:
        ids = []
        for k, inp in enumerate(good):
            ids.append(srv.submit_inputs(inp))
            if k < len(cases):
                mutate, status = cases[k]
                bad = gpu_ctx.inputs(jobs[k].camera, scene, jobs[k].lights, job_models(models, jobs[k]))
                mutate(bad)
                with pytest.raises(T3DError) as e:
                    srv.submit_inputs(bad)
                assert e.value.status == status
                with pytest.raises(T3DError) as e2:           # rendering it alone refuses it the same way
                    gpu_ctx.render_inputs(bad)
                assert e2.value.status == status
        for i in ids:
            srv.wait(i)
        shared = 0
        for k, i in enumerate(ids):
            r = srv.download(i)
            shared += r.info.batch_jobs > 1
            compare_frames(r, alone[k])
            srv.release(i)
        st = srv.stats()
        assert st.completed == len(good) and st.failed == 0 and shared > 0
        print("invalid jobs refused at submit; batches", st.batches, "jobs in shared batches", shared, flush=True)
    for m in models:
        if id(m.mesh) not in oracle._handles:
            oracle.register_mesh(m.mesh, m.mesh._handles["gpu"])
    compare_frames(alone[1], oracle.render(jobs[1].camera, scene, jobs[1].lights, job_models(models, jobs[1])))


def test_cell_lights_through_the_server(gpu_ctx, oracle):
    """A submitted job's LightVolume cell lights are copied with the job
    (use_lights deep-copied at submit) and render as Camera.Render does."""
    from tests.feature_scenes import cell_lights_scene
    from tetra3d_amd.serving import Server
    wl = cell_lights_scene(3)
    gpu_ctx.prepare(wl)
    ref = oracle.render_workload(wl)
    inp = gpu_ctx.inputs(wl.camera, wl.scene, wl.lights, wl.models)
    with Server(gpu_ctx, slots=1, max_batch_jobs=1) as srv:
        i = srv.submit_inputs(inp)
        del inp                     # the job holds its own copy of every descriptor
        srv.wait(i)
        compare_frames(srv.download(i), ref)
        srv.release(i)


def test_animation_and_volume_references(gpu_ctx):
    """A submitted job holds its animations and light volumes as it holds
    its meshes: freeing one returns BUSY until the job is released."""
    from tetra3d_amd.runtime import T3DError
    from tetra3d_amd.serving import Server
    from tetra3d_amd.scene import LightVolume
    wl = W.crowd_scene(count=4, bones=16, level=2, clips=2)
    vol = LightVolume("vol", (-50, -50, -80), (50, 50, 0), (10, 10, 10))
    vol.cells[:, :3] = 0.5
    lights = list(wl.lights) + [vol]
    gpu_ctx.prepare(wl)
    gpu_ctx.upload_volume(vol)
    anim = wl.models[0].animation
    with Server(gpu_ctx, slots=1, max_batch_jobs=1) as srv:
        i = srv.submit(wl.camera, wl.scene, lights, wl.models)
        with pytest.raises(T3DError) as e:
            gpu_ctx.free_animation(anim)
        assert e.value.status == A.T3D_ERR_BUSY
        rc = gpu_ctx.L.t3d_volume_free(gpu_ctx.ctx, vol.handle)
        assert rc == A.T3D_ERR_BUSY
        srv.wait(i)
        assert srv.download(i).survivors > 0
        srv.release(i)
    # released: both free now
    gpu_ctx.free_animation(anim)
    assert gpu_ctx.L.t3d_volume_free(gpu_ctx.ctx, vol.handle) == 0
    vol.handle = 0
    gpu_ctx.upload_animation(anim)      # leave the scene renderable for later tests


def test_failed_batch_build_retries_jobs_alone(pool):
    """When a batch of several jobs cannot be built (here: the context's HBM
    budget holds one job's frame but not the two-job batch's), its jobs run
    again one by one and both complete with their alone renders' results.
    auto_release hands each finished job's frame straight back, so the
    second job fits after the first."""
    from tetra3d_amd.runtime import Context
    from tetra3d_amd.serving import Server
    models, jobs, scene = pool
    big = sorted(jobs, key=lambda j: -sum(models[i].mesh.triangle_count for i in j.model_ids))[:2]
    with Context(0) as probe:                      # device memory of each job's frame alone
        for m in models:
            probe.upload_mesh(m.mesh)
        mesh_bytes = probe.memory_used
        frame_bytes, alone = [], []
        for j in big:
            f = probe.new_frame()
            before = probe.memory_used
            probe.render_async(probe.inputs(j.camera, scene, j.lights, job_models(models, j)), frame=f)
            probe.wait(frame=f)
            frame_bytes.append(probe.memory_used - before)
            alone.append(probe.pairs(frame=f))
    slack = 1 << 20
    assert min(frame_bytes) > 4 * slack, frame_bytes   # the two together cannot fit in one job's room
    with Context(0, hbm_budget=mesh_bytes + max(frame_bytes) + slack) as ctx:
        for m in models:
            ctx.upload_mesh(m.mesh)
        inputs = [ctx.inputs(j.camera, scene, j.lights, job_models(models, j)) for j in big]
        order = [0, 0, 1, 0, 1]      # the first occupies the one slot while the other four queue and batch
        with Server(ctx, slots=1, max_batch_jobs=8, auto_release=True) as srv:
            ids = [srv.submit_inputs(inputs[k]) for k in order]
            for k, i in zip(order, ids):
                srv.wait(i)
                info = srv.info(i)
                assert info.batch_jobs == 1
                assert info.pair_count == len(alone[k])
                assert info.survivors == sum(p[3] for p in alone[k])
            st = srv.stats()
            print("solo retry: frame bytes", frame_bytes, "batches", st.batches, "completed", st.completed,
                  "failed", st.failed, flush=True)
            # every job completed alone, and at least one batch of several failed to build and was split
            assert st.completed == len(order) and st.failed == 0 and st.batches > st.completed
