"""Pins the oracle's bone palettes from poses (Node.Transform down an
armature, node.go:572-608; boneInfluence = inverseBindMatrix·Transform(),
node.go:592-594) with closed forms that a swapped product order fails, a
float64 restatement on random trees, and the parent check.  Then the oracle's
render with poses equals its render with those palettes given directly.
CPU only."""
import math

import numpy as np
import pytest

from tetra3d_amd import _abi as A
from tests.feature_scenes import posed_scene, random_poses
from tests.parity import compare_frames

f32 = np.float32
RZ90 = np.array([[0, 1, 0, 0], [-1, 0, 0, 0], [0, 0, 1, 0], [0, 0, 0, 1]], dtype=np.float64)   # exact


def T(x, y, z):
    m = np.eye(4)
    m[3, :3] = (x, y, z)
    return m


def S(x, y, z):
    return np.diag([x, y, z, 1.0])


def poses_of(specs):
    """specs: (position, scale, rotation, parent, inverse_bind) per bone."""
    poses = (A.BonePose * max(1, len(specs)))()
    for b, (pos, sc, rot, parent, ib) in enumerate(specs):
        poses[b].position[:] = list(pos)
        poses[b].scale[:] = list(sc)
        poses[b].rotation = A.mat4(rot)
        poses[b].parent = parent
        poses[b].inverse_bind = A.mat4(ib)
    return poses


def test_scale_rotation_translation_order(oracle_lib):
    """Transform() = Scale·rotation·Translate: the translation row is the
    position as given, the rotation rows scaled per row."""
    pal = oracle_lib.bone_palettes(poses_of([((1, 2, 3), (2, 3, 4), RZ90, -1, np.eye(4))]))
    expect = S(2, 3, 4) @ RZ90 @ T(1, 2, 3)
    np.testing.assert_array_equal(pal[0], expect.astype(f32))
    assert not np.array_equal(expect, T(1, 2, 3) @ RZ90 @ S(2, 3, 4))
    assert not np.array_equal(expect, RZ90 @ S(2, 3, 4) @ T(1, 2, 3))


def test_inverse_bind_on_the_left(oracle_lib):
    ib = T(5, 0, 0)
    pal = oracle_lib.bone_palettes(poses_of([((0, 0, 0), (1, 1, 1), RZ90, -1, ib)]))
    np.testing.assert_array_equal(pal[0], (ib @ RZ90).astype(f32))
    assert not np.array_equal(ib @ RZ90, RZ90 @ ib)
    np.testing.assert_array_equal(pal[0][3], [0, 5, 0, 1])       # (5,0,0) rotated by the bone


def test_parent_and_armature_root_on_the_right(oracle_lib):
    """A child's Transform() is its local transform times its parent's; the
    root bones' times the armature parent's."""
    root = T(10, 20, 30)
    specs = [((0, 0, 0), (1, 1, 1), RZ90, -1, np.eye(4)),
             ((1, 0, 0), (1, 1, 1), np.eye(4), 0, np.eye(4)),
             ((0, 0, 2), (2, 2, 2), np.eye(4), 1, np.eye(4)),
             ((4, 0, 0), (1, 1, 1), np.eye(4), -1, np.eye(4))]
    pal = oracle_lib.bone_palettes(poses_of(specs), root)
    w0 = RZ90 @ root
    w1 = T(1, 0, 0) @ w0
    w2 = S(2, 2, 2) @ T(0, 0, 2) @ w1
    w3 = T(4, 0, 0) @ root
    for got, want in zip(pal, [w0, w1, w2, w3]):
        np.testing.assert_array_equal(got, want.astype(f32))
    np.testing.assert_array_equal(pal[1][3], [10, 21, 30, 1])    # (1,0,0) turned by the parent, then the root
    assert not np.array_equal(w1, w0 @ T(1, 0, 0))
    # without an armature parent the roots stay where they are
    pal = oracle_lib.bone_palettes(poses_of(specs))
    np.testing.assert_array_equal(pal[3], T(4, 0, 0).astype(f32))


@pytest.mark.parametrize("tree", [True, False], ids=["tree", "chain"])
def test_random_armatures_close_to_float64(oracle_lib, tree):
    rng = np.random.default_rng(11)
    n = 80
    poses = random_poses(rng, n, tree)
    root = np.asarray(T(*rng.uniform(-3, 3, 3)) @ np.diag([1.5, 1.5, 1.5, 1.0]), dtype=f32)
    pal = oracle_lib.bone_palettes(poses, root)
    world = []
    for b in range(n):
        p = poses[b]
        local = S(*p.scale) @ np.array(A.mat4_to_list(p.rotation)) @ T(*p.position)
        w = local @ (world[p.parent] if p.parent >= 0 else root.astype(np.float64))
        world.append(w)
        expect = np.array(A.mat4_to_list(p.inverse_bind)) @ w
        np.testing.assert_allclose(pal[b], expect, rtol=1e-4, atol=1e-5, err_msg="bone %d" % b)


def test_parent_must_be_an_earlier_bone(oracle_lib):
    for b, bad in [(0, 0), (2, 2), (2, 3), (1, -2)]:
        specs = [((0, 0, 0), (1, 1, 1), np.eye(4), -1 if k == 0 else k - 1, np.eye(4)) for k in range(4)]
        specs[b] = specs[b][:3] + (bad,) + specs[b][4:]
        assert oracle_lib.bone_palettes(poses_of(specs)) is None, (b, bad)


def test_render_with_poses_equals_render_with_palettes(oracle_lib):
    wl = posed_scene(0)
    o = oracle_lib.Oracle()
    try:
        posed = o.render_workload(wl)
        for m in wl.models:
            if m.bone_poses is not None:
                m.bone_influences = oracle_lib.bone_palettes(m.bone_poses, m.armature_parent)
                m.bone_poses = None
        given = o.render_workload(wl)
    finally:
        o.close()
    s = compare_frames(posed, given)
    assert s["survivors"] > 500
    assert not math.isnan(float(posed.depth.sum()))
