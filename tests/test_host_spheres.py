"""The frustum pass computes model bounding spheres (model.go:142-182) four
models at a time.  Host only: its spheres must equal the one-model path's and
the oracle's bit for bit, on random transforms and on the special cases of
Vector4.Unit (rows of length 1 and below 1e-8 are left as they are)."""
import ctypes as C

import numpy as np

from tetra3d_amd import _abi as A
from tetra3d_amd import runtime as R
from tetra3d_amd.scene import quat_to_matrix

f32 = np.float32


def _models(rng, n):
    t = np.zeros((n, 4, 4), dtype=f32)
    for i in range(n):
        q = rng.normal(size=4)
        q /= np.linalg.norm(q)
        rot = np.asarray(quat_to_matrix(*q.astype(f32)), dtype=f32)[:3, :3]
        s = rng.uniform(0.05, 20, 3) * rng.choice([-1.0, 1.0], 3)
        t[i, :3, :3] = rot * s[:, None].astype(f32)
        t[i, 3, :3] = rng.uniform(-500, 500, 3)
        t[i, 3, 3] = 1
        if i % 7 == 0:
            t[i, :3, 3] = rng.normal(size=3)           # a non-affine w column
    t[1, :3, :3] = np.eye(3, dtype=f32)                # unit rows: left as they are
    t[2, 0, :3] = 0                                    # a zero row
    t[3, 1, :3] = [1e-9, 0, 0]                         # a row shorter than 1e-8
    t[4, :3, :3] = np.diag([1, -1, 1]).astype(f32)
    lo = rng.uniform(-5, 0, (n, 3))
    hi = lo + rng.uniform(0, 10, (n, 3))
    dims = np.concatenate([lo, hi], axis=1).astype(f32)
    dims[5] = [-1, -1, -1, 1, 1, 1]                    # centred dims: a zero centre
    skinned = (rng.random(n) < 0.3).astype(np.uint8)
    return t, dims, skinned


def test_batched_spheres_match_scalar_and_oracle(oracle_lib):
    rng = np.random.default_rng(3)
    n = 4099                                           # not a multiple of four: the tail batch
    t, dims, skinned = _models(rng, n)
    batched, scalar = R.model_spheres(t, dims, skinned)
    np.testing.assert_array_equal(batched.view(np.uint32), scalar.view(np.uint32))
    L = oracle_lib.lib()
    pos, rad = (C.c_float * 3)(), C.c_float()
    for i in range(n):
        L.t3do_model_frustum_sphere(C.byref(A.mat4(t[i])), int(skinned[i]), (C.c_float * 3)(*dims[i, :3]),
                                    (C.c_float * 3)(*dims[i, 3:]), pos, C.byref(rad))
        expect = np.array([pos[0], pos[1], pos[2], rad.value], dtype=f32)
        np.testing.assert_array_equal(batched[i].view(np.uint32), expect.view(np.uint32), err_msg="model %d" % i)
