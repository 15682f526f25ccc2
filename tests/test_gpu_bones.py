"""Bone palettes computed on the device from poses (node.go:572-608) against
the oracle, bit for bit: both execution forms, armatures in and past LDS
and past the level-by-level walk (5,000 bones), a palette shared by two
render pairs, frames that change only the pose (the captured chain replayed
with new poses), equality with host palettes, the parent check, animations
sampled on the device (AnimationPlayer.updateValues, animation.go:674-720)
and their refused inputs, and posed and animated jobs in the serving mix."""
import numpy as np
import pytest

from tetra3d_amd import _abi as A
from tetra3d_amd import workloads as W
from tetra3d_amd.runtime import T3DError
from tests.feature_scenes import feature_scene, huge_armature_scene, posed_scene
from tests.parity import compare_frames

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("mode", [A.EXEC_BATCHED, A.EXEC_PLAIN], ids=["batched", "plain"])
@pytest.mark.parametrize("seed", [0, 1])
def test_posed_matches_oracle(gpu_ctx, oracle, seed, mode):
    wl = posed_scene(seed)
    print("posed scene", seed, "mode", mode, flush=True)
    s = compare_frames(gpu_ctx.render_workload(wl, exec_mode=mode), oracle.render_workload(wl))
    assert s["survivors"] > 500


@pytest.mark.parametrize("mode", [A.EXEC_BATCHED, A.EXEC_PLAIN], ids=["batched", "plain"])
def test_armature_past_level_walk(gpu_ctx, oracle, mode):
    wl = huge_armature_scene(9)
    print("5000-bone armature, mode", mode, flush=True)
    s = compare_frames(gpu_ctx.render_workload(wl, exec_mode=mode), oracle.render_workload(wl))
    assert s["survivors"] > 100


def test_pose_changes_between_frames(gpu_ctx, oracle):
    """Same launch shape every frame, so the batched form replays its graph;
    only the poses and the armature parent move."""
    wl = posed_scene(2)
    rng = np.random.default_rng(5)
    for frame in range(4):
        print("pose frame", frame, flush=True)
        for m in wl.models:
            if m.bone_poses is None:
                continue
            for b in range(0, len(m.bone_poses), 3):
                axis = W.unit_vectors(W.SplitMix64(int(rng.integers(1 << 30))), 1)[0]
                m.bone_poses[b].rotation = A.mat4(W.axis_angle_matrix(axis, float(rng.uniform(0, 0.4))))
            if m.armature_parent is not None:
                m.armature_parent[3, 1] = np.float32(0.25 * frame)
        compare_frames(gpu_ctx.render_workload(wl), oracle.render_workload(wl))


def test_device_palettes_equal_host_palettes(gpu_ctx, oracle_lib):
    wl = posed_scene(3)
    posed = gpu_ctx.render_workload(wl)
    for m in wl.models:
        if m.bone_poses is not None:
            m.bone_influences = oracle_lib.bone_palettes(m.bone_poses, m.armature_parent)
            m.bone_poses = None
    given = gpu_ctx.render_workload(wl)
    compare_frames(posed, given)


def test_parent_after_child_is_refused(gpu_ctx, oracle):
    wl = posed_scene(4)
    gpu_ctx.prepare(wl)
    big = wl.models[1].bone_poses
    for bad in (7, 200, -2):
        old = big[7].parent
        big[7].parent = bad
        try:
            with pytest.raises(T3DError) as e:
                gpu_ctx.render_workload(wl)
            assert e.value.status == A.T3D_ERR_INVALID_ARGUMENT
        finally:
            big[7].parent = old
    compare_frames(gpu_ctx.render_workload(wl), oracle.render_workload(wl))


def animated_scene(seed):
    """posed_scene with every posed armature playing an animation with fewer
    channels than it has bones (so the last bones keep their poses)."""
    wl = posed_scene(seed)
    rng = W.SplitMix64(seed + 100)
    for i, m in enumerate(wl.models):
        if m.bone_poses is not None:
            m.animation = W.random_animation(rng.child(i), max(1, len(m.bone_poses) - 3), keys=9 + i)
            m.animation_time = 0.3 + 0.2 * i
            # a one-keyframe track (always that keyframe) and a zero quaternion (the pose's rotation stays)
            m.animation.add_track(0, "scale", [0.5], [[1.25, 0.75, 1.0]])
            if len(m.animation.channels) > 2:
                m.animation.add_track(2, "rotation", [0.0, 1.0], [[0, 0, 0, 0], [0, 0, 0, 0]])
    return wl


@pytest.mark.parametrize("mode", [A.EXEC_BATCHED, A.EXEC_PLAIN], ids=["batched", "plain"])
def test_animated_matches_oracle(gpu_ctx, oracle, mode):
    """Playheads before, inside, on and after the keyframes, frame after
    frame (the batched form replays its graph with new playheads)."""
    wl = animated_scene(10)
    for frame, t in enumerate([-0.5, 0.0, 0.125, 0.37, 0.5, 0.99, 1.0, 2.0]):
        for i, m in enumerate(wl.models):
            if m.animation is not None:
                m.animation_time = t + 0.01 * i * (frame % 2)
        print("animated frame", frame, "t", t, "mode", mode, flush=True)
        compare_frames(gpu_ctx.render_workload(wl, exec_mode=mode), oracle.render_workload(wl))


def test_animation_inputs_refused(gpu_ctx, oracle):
    from tetra3d_amd.scene import Animation
    bad = []
    an = Animation("descending", 1)
    an.add_track(0, "position", [0.0, 1.0, 0.5], [[0, 0, 0]] * 3)
    bad.append(an)
    an = Animation("nan_time", 1)
    an.add_track(0, "rotation", [0.0, float("nan")], [[0, 0, 0, 1]] * 2)
    bad.append(an)
    an = Animation("interpolation", 1)
    an.add_track(0, "scale", [0.0, 1.0], [[1, 1, 1]] * 2, interpolation=7)
    bad.append(an)
    for an in bad:
        print("refused upload", an.name, flush=True)
        with pytest.raises(T3DError) as e:
            gpu_ctx.upload_animation(an)
        assert e.value.status == A.T3D_ERR_INVALID_ARGUMENT, an.name
    wl = animated_scene(11)
    gpu_ctx.prepare(wl)
    m = wl.models[0]
    m.animation_time = float("nan")
    with pytest.raises(T3DError) as e:
        gpu_ctx.render_workload(wl)
    assert e.value.status == A.T3D_ERR_INVALID_ARGUMENT
    m.animation_time = 0.5
    poses, m.bone_poses = m.bone_poses, None          # an animation needs poses
    m.bone_influences = np.tile(np.eye(4, dtype=np.float32), (12, 1, 1))
    with pytest.raises(T3DError) as e:
        gpu_ctx.render_workload(wl)
    assert e.value.status == A.T3D_ERR_INVALID_ARGUMENT
    m.bone_poses = poses
    gone = m.animation
    gpu_ctx.free_animation(gone)
    gone.handle = 12345678                            # a handle the context does not know
    with pytest.raises(T3DError) as e:
        gpu_ctx.render_workload(wl)
    assert e.value.status == A.T3D_ERR_NOT_FOUND
    gone.handle = 0
    compare_frames(gpu_ctx.render_workload(wl), oracle.render_workload(wl))


def test_posed_jobs_same_alone_and_in_mix(gpu_ctx):
    from tetra3d_amd.serving import Server
    scenes = [posed_scene(5), feature_scene(6), animated_scene(7), posed_scene(8)]
    alone = [gpu_ctx.render_workload(wl) for wl in scenes]
    inputs = [gpu_ctx.inputs(wl.camera, wl.scene, wl.lights, wl.models) for wl in scenes]
    with Server(gpu_ctx, slots=2, max_batch_jobs=4) as srv:
        ids = [srv.submit_inputs(inp) for inp in inputs]
        for i in ids:
            srv.wait(i)
        shared = 0
        for k, i in enumerate(ids):
            r = srv.download(i)
            shared += r.info.batch_jobs > 1
            compare_frames(r, alone[k])
            srv.release(i)
    print("posed jobs in shared batches", shared, flush=True)


def test_submitted_poses_are_copied(gpu_ctx):
    """The server copies a job's poses, armature parent and playhead at
    submit: changing them afterwards leaves the job's result as it was."""
    from tetra3d_amd.serving import Server
    wl = animated_scene(12)
    before = gpu_ctx.render_workload(wl)
    inp = gpu_ctx.inputs(wl.camera, wl.scene, wl.lights, wl.models)
    with Server(gpu_ctx, slots=1, max_batch_jobs=2) as srv:
        job = srv.submit_inputs(inp)
        for m in wl.models:                       # scribble over what the descriptor pointed to
            if m.bone_poses is not None:
                for b in range(len(m.bone_poses)):
                    m.bone_poses[b].position[:] = [100.0, 100.0, 100.0]
        for i in range(inp.render.model_count):
            md = inp.render.models[i]
            if md.armature_parent:
                md.armature_parent[0].m[3][0] = 500.0
            md.animation_time = 7.0
        srv.wait(job)
        r = srv.download(job)
        compare_frames(r, before)
        srv.release(job)
