"""Pins the oracle's animation sampling (AnimationTrack.ValueAsVector /
ValueAsQuaternion, animation.go:63-160; forceUpdate, animation.go:902-983)
and the math under it: Go's float64 math.Sin and math.Atan2 as restated
(checked against the C library to within an ulp, and on Go's documented
special cases), Quaternion.Slerp against its closed form and float64
evaluation, ToMatrix4 against axis-angle rotations, and keyframe search
edge cases with exact dyadic values.  CPU only."""
import ctypes as C
import math

import numpy as np
import pytest

from tetra3d_amd import _abi as A
from tetra3d_amd import workloads as W
from tetra3d_amd.scene import Animation
from tests.feature_scenes import posed_scene
from tests.parity import compare_frames

f32 = np.float32
FP = C.POINTER(C.c_float)


@pytest.fixture(scope="module")
def L(oracle_lib):
    return oracle_lib.lib()


def ulps(a, b):
    ia = np.array([a, b], dtype=np.float64).view(np.int64)
    ia = np.where(ia < 0, np.int64(-0x8000000000000000) - ia, ia)
    return abs(int(ia[0]) - int(ia[1]))


def test_go_sin_against_libm(L):
    rng = np.random.default_rng(1)
    xs = np.concatenate([rng.uniform(-20, 20, 100000), rng.uniform(0, math.pi / 2, 100000)])
    worst, f32_diff = 0, 0
    for x in xs:
        g, m = L.t3do_go_sin(float(x)), math.sin(float(x))
        worst = max(worst, ulps(g, m))
        f32_diff += f32(g) != f32(m)
    assert worst <= 2, worst
    assert f32_diff <= 2, f32_diff          # float32 results (what math32.Sin returns) agree
    for x in np.linspace(-math.pi, math.pi, 4097):   # near the zeros at +-pi: absolute error
        assert abs(L.t3do_go_sin(float(x)) - math.sin(float(x))) <= 2.3e-16
    # Go's special cases: sin(+-0) = +-0, sin(NaN) = NaN, sin(+-Inf) = NaN
    assert math.copysign(1, L.t3do_go_sin(-0.0)) == -1 and L.t3do_go_sin(0.0) == 0
    assert math.isnan(L.t3do_go_sin(float("nan"))) and math.isnan(L.t3do_go_sin(float("inf")))
    # exactness where the octant reduction is trivial
    assert L.t3do_go_sin(1e-300) == 1e-300


def test_go_atan2_against_libm_and_special_cases(L):
    rng = np.random.default_rng(2)
    worst = 0
    for _ in range(60000):
        y, x = rng.normal(size=2) * 10.0 ** rng.integers(-3, 4, 2)
        worst = max(worst, ulps(L.t3do_go_atan2(y, x), math.atan2(y, x)))
    assert worst <= 2, worst
    for y in np.linspace(0.001, 1, 3000):          # Slerp's quadrant: sinOmega > 0, cosOmega >= 0
        x = math.sqrt(max(0.0, 1 - y * y))
        assert ulps(L.t3do_go_atan2(y, x), math.atan2(y, x)) <= 1
    inf, nan = float("inf"), float("nan")
    cases = [((nan, 1.0), nan), ((1.0, nan), nan), ((0.0, 1.0), 0.0), ((-0.0, 1.0), -0.0),
             ((0.0, -1.0), math.pi), ((-0.0, -1.0), -math.pi), ((0.0, -0.0), math.pi), ((1.0, 0.0), math.pi / 2),
             ((-1.0, -0.0), -math.pi / 2), ((inf, inf), math.pi / 4), ((-inf, -inf), -3 * math.pi / 4),
             ((1.0, inf), 0.0), ((1.0, -inf), math.pi), ((inf, 1.0), math.pi / 2)]
    for (y, x), want in cases:
        got = L.t3do_go_atan2(y, x)
        if math.isnan(want):
            assert math.isnan(got)
        else:
            assert got == want and math.copysign(1, got) == math.copysign(1, want), (y, x, got, want)


def slerp(L, a, b, t):
    out = (C.c_float * 4)()
    L.t3do_quat_slerp((C.c_float * 4)(*a), (C.c_float * 4)(*b), t, out)
    return np.array(out[:], dtype=f32)


def to_mat(L, q):
    out = A.Mat4()
    L.t3do_quat_to_mat4((C.c_float * 4)(*q), C.byref(out))
    return np.array(A.mat4_to_list(out), dtype=f32)


def test_slerp_closed_forms(L):
    ident = (0, 0, 0, 1)
    z90 = (0, 0, math.sin(math.pi / 4), math.cos(math.pi / 4))
    for t in (0.25, 0.5, 0.75):
        got = slerp(L, ident, z90, t)
        want = [0, 0, math.sin(t * math.pi / 4), math.cos(t * math.pi / 4)]
        np.testing.assert_allclose(got, want, atol=2e-7)
    # the shortest arc: the far copy of the end quaternion gives the same bits
    np.testing.assert_array_equal(slerp(L, ident, z90, 0.3), slerp(L, ident, [-v for v in z90], 0.3))
    # nearly equal rotations take the linear branch: (1 - t) a + t b exactly
    a, b = (0, 0, 0.001, 0.9999995), (0, 0, 0.0012, 0.99999928)
    k0, k1 = f32(1) - f32(0.4), f32(0.4)
    want = [f32(a[i]) * k0 + f32(b[i]) * k1 for i in range(4)]
    np.testing.assert_array_equal(slerp(L, a, b, 0.4), np.array(want, dtype=f32))
    # float64 slerp on random pairs
    rng = np.random.default_rng(3)
    for _ in range(2000):
        p, q = rng.normal(size=(2, 4))
        p, q = p / np.linalg.norm(p), q / np.linalg.norm(q)
        t = rng.random()
        c = float(p @ q)
        if c < 0:
            q, c = -q, -c
        if c > 0.9999:
            continue
        om = math.acos(c)
        want = (math.sin((1 - t) * om) * p + math.sin(t * om) * q) / math.sin(om)
        np.testing.assert_allclose(slerp(L, p.astype(f32), q.astype(f32), t), want, atol=3e-6)


def test_quaternion_matrix(L):
    for axis, ang in [((0, 0, 1), math.pi / 2), ((1, 0, 0), 0.3), ((0.6, 0.0, 0.8), 2.0), ((0, 1, 0), -1.1)]:
        s = math.sin(ang / 2)
        q = (axis[0] * s, axis[1] * s, axis[2] * s, math.cos(ang / 2))
        np.testing.assert_allclose(to_mat(L, q), np.asarray(W.axis_angle_matrix(axis, ang)), atol=2e-6)
    np.testing.assert_array_equal(to_mat(L, (0, 0, 0, 1)), np.eye(4, dtype=f32))
    np.testing.assert_array_equal(to_mat(L, (0, 0, 0, 0)), np.zeros((4, 4), dtype=f32))


def _poses(n):
    poses = (A.BonePose * n)()
    for b in range(n):
        poses[b].position[:] = [7.0, 8.0, 9.0]
        poses[b].scale[:] = [1.0, 1.0, 1.0]
        poses[b].parent = b - 1
        poses[b].rotation = A.mat4(np.eye(4))
        poses[b].inverse_bind = A.mat4(np.eye(4))
    return poses


def test_keyframe_search_and_interpolation(oracle_lib):
    an = Animation("probe", 4)
    an.add_track(0, "position", [0.0, 1.0, 2.0, 2.0, 3.0], [[0, 0, 0], [10, 20, 40], [20, 0, 0], [30, 0, 0],
                                                          [50, 0, 0]])
    an.add_track(1, "position", [0.0, 1.0], [[0, 0, 0], [8, 8, 8]], A.INTERPOLATION_CONSTANT)
    an.add_track(2, "scale", [0.0, 1.0], [[1, 1, 1], [3, 5, 9]], A.INTERPOLATION_CUBIC)   # samples as linear
    z90 = (0, 0, math.sin(math.pi / 4), math.cos(math.pi / 4))
    an.add_track(3, "rotation", [0.0, 1.0], [(0, 0, 0, 1), z90])
    poses = _poses(5)

    def at(t):
        return oracle_lib.animate_poses(an, t, poses)

    p = at(-1.0)                       # before the first keyframe: the first value
    assert p[0].position[:] == [0, 0, 0] and p[1].position[:] == [0, 0, 0]
    p = at(0.25)                       # linear between dyadic keys: exact
    assert p[0].position[:] == [2.5, 5.0, 10.0]
    assert p[1].position[:] == [0, 0, 0]          # constant: the earlier keyframe
    assert p[2].scale[:] == [1.5, 2.0, 3.0]       # cubic is linear
    p = at(1.0)                        # exactly on a keyframe: its value, no interpolation
    assert p[0].position[:] == [10, 20, 40] and p[1].position[:] == [8, 8, 8]
    p = at(2.0)                        # two keyframes at the same time: the first of them
    assert p[0].position[:] == [20, 0, 0]
    p = at(2.5)                        # after the duplicate: from the last of them
    assert p[0].position[:] == [40, 0, 0]
    p = at(9.0)                        # after the last keyframe: the last value
    assert p[0].position[:] == [50, 0, 0]
    np.testing.assert_allclose(np.array(A.mat4_to_list(at(0.5)[3].rotation)),
                               np.asarray(W.axis_angle_matrix((0, 0, 1), math.pi / 4)), atol=2e-6)
    p = at(0.5)                        # tracks absent: the pose stays; bones past the channels stay
    assert p[1].scale[:] == [1, 1, 1] and p[4].position[:] == [7, 8, 9]
    assert p[0].scale[:] == [1, 1, 1] and np.array_equal(np.array(A.mat4_to_list(p[0].rotation)), np.eye(4))


def test_single_keyframe_track(oracle_lib):
    an = Animation("single", 1)
    an.add_track(0, "position", [0.5], [[1, 2, 3]])
    for t in (-1.0, 0.5, 0.75, 9.0):           # before, on and after its one keyframe
        assert oracle_lib.animate_poses(an, t, _poses(1))[0].position[:] == [1, 2, 3]


def test_zero_quaternion_keeps_rotation(oracle_lib):
    an = Animation("zero", 1)
    an.add_track(0, "rotation", [0.0, 1.0], [(0, 0, 0, 0), (0, 0, 0, 0)])
    poses = _poses(1)
    poses[0].rotation = A.mat4(W.axis_angle_matrix((1, 0, 0), 0.5))
    p = oracle_lib.animate_poses(an, 0.5, poses)
    assert A.mat4_to_list(p[0].rotation) == A.mat4_to_list(poses[0].rotation)


def test_render_with_animation_equals_render_with_sampled_poses(oracle_lib):
    wl = posed_scene(0)
    rng = W.SplitMix64(44)
    for i, m in enumerate(wl.models):
        if m.bone_poses is not None:
            m.animation = W.random_animation(rng.child(i), len(m.bone_poses) - 2)
            m.animation_time = 0.37 + 0.1 * i
    o = oracle_lib.Oracle()
    try:
        animated = o.render_workload(wl)
        for m in wl.models:
            if m.animation is not None:
                m.bone_poses = oracle_lib.animate_poses(m.animation, m.animation_time, m.bone_poses)
                m.animation = None
        sampled = o.render_workload(wl)
    finally:
        o.close()
    compare_frames(animated, sampled)
    assert animated.survivors > 500
