"""Pins the oracle's math restatement (matrix.go, vector.go, camera.go
projection/frustum, model.go frustum sphere and invertedCamPos) against the
reference's own test (matrix_test.go:36-55), closed forms and invariants.
CPU only."""
import ctypes as C
import math

import numpy as np
import pytest

from tetra3d_amd import _abi as A
from tetra3d_amd.scene import Camera, mat_mult, mat_scale, mat_translate

f32 = np.float32
FP = C.POINTER(C.c_float)


@pytest.fixture(scope="module")
def L(oracle_lib):
    return oracle_lib.lib()


def m4(a):
    return A.mat4(np.asarray(a, dtype=f32))


def arr(m):
    return np.array(A.mat4_to_list(m), dtype=f32)


def mult(L, a, b):
    out = A.Mat4()
    L.t3do_mat4_mult(C.byref(m4(a)), C.byref(m4(b)), C.byref(out))
    return arr(out)


def inverted(L, a):
    out = A.Mat4()
    L.t3do_mat4_inverted(C.byref(m4(a)), C.byref(out))
    return arr(out)


def new_rotate(x, y, z, angle):
    """NewMatrix4Rotate (matrix.go:101-128) in float32 (test input construction)."""
    x, y, z, angle = f32(x), f32(y), f32(z), f32(angle)
    if x == 0 and y == 0 and z == 0:
        y = f32(1)
    l = f32(math.sqrt(float(x * x + y * y + z * z)))
    if not (l < f32(1e-8) or l == f32(1)):
        x, y, z = x / l, y / l, z / l
    s, c = f32(math.sin(float(angle))), f32(math.cos(float(angle)))
    m = f32(1) - c
    r = np.eye(4, dtype=f32)
    r[0, 0] = m * x * x + c
    r[0, 1] = m * x * y + z * s
    r[0, 2] = m * z * x - y * s
    r[1, 0] = m * x * y - z * s
    r[1, 1] = m * y * y + c
    r[1, 2] = m * y * z + x * s
    r[2, 0] = m * z * x + y * s
    r[2, 1] = m * y * z - x * s
    r[2, 2] = m * z * z + c
    return r


def is_identity_go(m):
    """Matrix4.IsIdentity (matrix.go:511-529): |a-b| > 1e-4 fails; NaN compares false."""
    with np.errstate(invalid="ignore"):
        return not np.any(np.abs(m - np.eye(4, dtype=f32)) > f32(1e-4))


def test_reference_inversion_test_matrices(L):
    """matrix_test.go:36-55, same four matrices and the same identity check."""
    mats = [
        new_rotate(0, 1, 0, 0.1),
        mat_translate(-10, 0.1, 3232.1976),
        mat_scale(10, 0.1, -0.45),
        mult(L, mult(L, mat_translate(-1, -1, -1), new_rotate(1, 0, 0.1, 0.334)), mat_scale(10, 1, 0)),
    ]
    with np.errstate(all="ignore"):
        for i, m in enumerate(mats):
            assert is_identity_go(mult(L, m, inverted(L, m))), i
    # the fourth matrix is singular (scale z = 0): the reference's check passes
    # only because NaN never compares greater than the epsilon
    with np.errstate(all="ignore"):
        assert np.isnan(mult(L, mats[3], inverted(L, mats[3]))).any()


def test_mult_identity_and_float64_agreement(L):
    rng = np.random.default_rng(3)
    I = np.eye(4, dtype=f32)
    for _ in range(50):
        a = rng.uniform(-10, 10, (4, 4)).astype(f32)
        b = rng.uniform(-10, 10, (4, 4)).astype(f32)
        assert np.array_equal(mult(L, a, I), a)
        assert np.array_equal(mult(L, I, a), a)
        ref = a.astype(np.float64) @ b.astype(np.float64)
        np.testing.assert_allclose(mult(L, a, b), ref, rtol=1e-5, atol=1e-4)


def test_mult_accumulation_order(L):
    """((a0*b0 + a1*b1) + a2*b2) + a3*b3: with 1e8, 1, -1e8, 1 the left-to-right
    float32 sum is 1 (the 1 is absorbed by 1e8), any other grouping gives 0 or 2."""
    a = np.zeros((4, 4), dtype=f32)
    a[0] = [1e8, 1, -1e8, 1]
    b = np.eye(4, dtype=f32)
    b[:, 0] = 1
    r = mult(L, a, b)
    assert r[0, 0] == f32(1)


def test_inverse_invariant_random_affine(L):
    rng = np.random.default_rng(4)
    for _ in range(50):
        m = np.eye(4, dtype=f32)
        m[:3, :3] = rng.uniform(-2, 2, (3, 3)) + 3 * np.eye(3)
        m[3, :3] = rng.uniform(-50, 50, 3)
        p = mult(L, m, inverted(L, m))
        np.testing.assert_allclose(p, np.eye(4), atol=2e-4)


def test_decompose_closed_form(L):
    """Decompose (matrix.go:181-196) of S·R·T: position = T exactly, scale and
    rotation recovered."""
    rot = new_rotate(0.3, 1, -0.2, 0.7)
    m = mat_mult(mat_mult(mat_scale(2, 3, 0.5), rot), mat_translate(5, -7, 11))
    pos, scl, r = (C.c_float * 3)(), (C.c_float * 3)(), A.Mat4()
    L.t3do_mat4_decompose(C.byref(m4(m)), pos, scl, C.byref(r))
    assert list(pos) == [f32(5), f32(-7), f32(11)]
    np.testing.assert_allclose(list(scl), [2, 3, 0.5], rtol=1e-6)
    np.testing.assert_allclose(arr(r)[:3, :3], rot[:3, :3], atol=1e-6)


def test_unit_quirks(L):
    """Vector3.Unit (vector.go:258-266): unchanged below 1e-8 or at exactly 1."""
    out = (C.c_float * 3)()
    for v in ([0, 0, 0], [1e-9, 0, 0], [1, 0, 0], [0.6, 0.8, 0]):
        L.t3do_vec3_unit((C.c_float * 3)(*v), out)
        assert list(out) == [float(f32(x)) for x in v]
    L.t3do_vec3_unit((C.c_float * 3)(3, 4, 0), out)
    assert list(out) == [float(f32(0.6)), float(f32(0.8)), 0.0]


def camera_desc(w=640, h=360, fov=75, near=0.1, far=100, pos=(0, 0, 0)):
    cam = Camera("c", w, h)
    cam.fov, cam.near, cam.far = f32(fov), f32(near), f32(far)
    cam.set_local_position(*pos)
    return cam.desc()


def test_projection_closed_forms(L):
    """VP on a camera at the origin looking down -Z: a view-space point
    (x, 0, -d) gives X/W = x(f-n)/(2 f d t aspect) and Z+1 = d(f+n)/(f-n)."""
    d = camera_desc()
    vp = A.Mat4()
    L.t3do_vp_matrix(C.byref(d), C.byref(vp))
    n, f = 0.1, 100.0
    t = math.tan(75 * math.pi / 360)
    aspect = 640 / 360
    for x, dist in [(1.0, 4.0), (-2.5, 10.0), (0.3, 0.5)]:
        out = (C.c_float * 4)()
        L.t3do_mat4_mult_vec_w(C.byref(vp), (C.c_float * 3)(x, 0, -dist), out)
        X, Y, Z, W = (float(v) for v in out)
        assert Y == 0
        np.testing.assert_allclose(X / W, x * (f - n) / (2 * f * dist * t * aspect), rtol=2e-6)
        np.testing.assert_allclose(Z + 1, dist * (f + n) / (f - n), rtol=2e-6)
    # the frustum's right edge, x = d t aspect, maps to (f-n)/(2f), not 1
    out = (C.c_float * 4)()
    L.t3do_mat4_mult_vec_w(C.byref(vp), (C.c_float * 3)(4 * t * aspect, 0, -4), out)
    np.testing.assert_allclose(out[0] / out[3], (f - n) / (2 * f), rtol=2e-6)


def test_sphere_in_frustum(L):
    d = camera_desc(pos=(0, 0, 4), far=100)
    F = lambda p, r: L.t3do_sphere_in_frustum(C.byref(d), (C.c_float * 3)(*p), f32(r))
    assert F((0, 0, 0), 1)            # straight ahead
    assert not F((0, 0, 10), 1)       # behind the camera
    assert not F((0, 0, -120), 5)     # beyond far + radius
    assert F((0, 0, -100), 5)         # straddles the far plane
    assert not F((40, 0, 0), 1)       # far outside the side planes
    assert F((6.5, 0, 0), 1)          # inside the horizontal half-angle (aspect 16:9)
    assert not F((0, 6.5, 0), 1)      # outside the vertical one


def test_model_frustum_sphere_closed_form(L):
    """model.go:142-182: centre = position + rotated dims centre * scale,
    radius = |scaled (max-min)| / 2; skinned models ignore the dims centre."""
    rot = new_rotate(0, 0, 1, math.pi / 2)
    m = mat_mult(mat_mult(mat_scale(2, 2, 2), rot), mat_translate(10, 0, 0))
    pos, rad = (C.c_float * 3)(), C.c_float()
    L.t3do_model_frustum_sphere(C.byref(m4(m)), 0, (C.c_float * 3)(0, 0, 0), (C.c_float * 3)(2, 0, 0), pos, C.byref(rad))
    np.testing.assert_allclose(list(pos), [10, 2, 0], atol=1e-5)   # centre (1,0,0) rotated 90° about z, x2
    np.testing.assert_allclose(rad.value, 2.0, rtol=1e-6)
    L.t3do_model_frustum_sphere(C.byref(m4(m)), 1, (C.c_float * 3)(0, 0, 0), (C.c_float * 3)(2, 0, 0), pos, C.byref(rad))
    np.testing.assert_allclose(list(pos), [10, 0, 0], atol=1e-6)


def test_inverted_cam_pos_quirk(L):
    """invertedCamPos (model.go:629-630) for M = Scale(s)·Translate(t) is
    camPos - t, not the true local position (camPos - t)/s."""
    m = mat_mult(mat_scale(2, 2, 2), mat_translate(1, 2, 3))
    out = (C.c_float * 3)()
    L.t3do_inverted_cam_pos(C.byref(m4(m)), (C.c_float * 3)(5, 5, 5), out)
    np.testing.assert_allclose(list(out), [4, 3, 2], atol=1e-5)
    m = mat_translate(1, 2, 3)
    L.t3do_inverted_cam_pos(C.byref(m4(m)), (C.c_float * 3)(5, 5, 5), out)
    assert list(out) == [4, 3, 2]


def test_bucket_bins_formula(L):
    """Sort's bin = int(clamp(((d-min)/(max-min))*512, 0, 511)), range 0 -> +0.001."""
    B = lambda d, lo, hi, mode=0: L.t3do_bucket_bin(f32(d), f32(lo), f32(hi), mode)
    assert B(0, 0, 1) == 0
    assert B(1, 0, 1) == 511          # 512 clamps to 511
    assert B(0.5, 0, 1) == 256
    assert B(0.25, 0, 1) == 128
    assert B(5, 5, 5) == 0            # range 0 -> 0.001, (5-5)/0.001 = 0
    assert B(0.7, 0, 1, A.SORT_NONE) == 0


def _brute_order(depths, mode):
    lo, hi = min(depths), max(depths)
    rd = hi - lo if hi != lo else 0.001
    bins = [0 if mode == A.SORT_NONE else min(511, max(0, int((d - lo) / rd * 512))) for d in depths]
    if mode == A.SORT_BACK_TO_FRONT:
        return sorted(range(len(depths)), key=lambda i: (-bins[i], i))
    return sorted(range(len(depths)), key=lambda i: (bins[i], i))


@pytest.mark.parametrize("mode", [A.SORT_BACK_TO_FRONT, A.SORT_FRONT_TO_BACK, A.SORT_NONE])
def test_bucket_sort_brute_force(L, mode):
    """sortingTriangleBucket.Sort + ForEach (sortingTriangles.go:38-125) against a
    brute-force stable sort by bin, on depths kept away from bin edges."""
    rng = np.random.default_rng(5)
    for n in (1, 2, 7, 100, 2000):
        bins = rng.integers(0, 512, n)
        depths = ((bins + 0.5) / 512.0).astype(np.float64)
        depths[0], depths[-1] = 0.0, 1.0            # pin min and max
        d32 = depths.astype(f32)
        order = (C.c_uint32 * n)()
        L.t3do_bucket_sort(d32.ctypes.data_as(FP), n, f32(d32.min()), f32(d32.max()), mode, order)
        assert list(order) == _brute_order(list(depths), mode)
        assert sorted(order) == list(range(n))      # a permutation


def test_go_stable_sort_is_stable_for_strict_orders(L):
    """The restated sort.SliceStable equals a stable sort for strict weak orders."""
    rng = np.random.default_rng(6)
    for n in (0, 1, 5, 19, 20, 21, 40, 41, 100, 333):
        keys = rng.integers(0, 7, n)
        table = (keys[:, None] < keys[None, :]).astype(np.uint8).ravel() if n else np.zeros(1, np.uint8)
        items = (C.c_int32 * max(n, 1))(*range(n))
        L.t3do_go_stable_sort_table(items, n, table.ctypes.data_as(C.POINTER(C.c_uint8)))
        assert list(items)[:n] == sorted(range(n), key=lambda i: (keys[i], i))
