"""The oracle renders every seeded feature scene, deterministically, and the
scenes exercise what they are meant to (the GPU parity test relies on both)."""
import numpy as np

from tetra3d_amd import _abi as A
from tests.feature_scenes import feature_scene


def test_feature_scenes_render_and_repeat(oracle):
    sort_modes, transparency, lv_modes = set(), set(), set()
    skinned = shape_keys = groups = ortho = batched_static = blend_off = offset = hidden = normals = 0
    billboards = unbillboarded = batchers = 0
    for seed in range(40):
        wl = feature_scene(seed)
        a = oracle.render_workload(wl)
        b = oracle.render_workload(wl)
        assert [p[:4] for p in a.pairs] == [p[:4] for p in b.pairs]
        np.testing.assert_array_equal(a.color.view(np.uint32), b.color.view(np.uint32))
        np.testing.assert_array_equal(a.triangle_ids, b.triangle_ids)
        ortho += not wl.camera.perspective
        normals += wl.camera.render_normals
        if wl.camera.render_normals:
            assert a.normals is not None and a.normals.shape == a.color.shape
            np.testing.assert_array_equal(a.normals.view(np.uint32), b.normals.view(np.uint32))
        offset += sum(bool(np.any(getattr(l, "offset", np.zeros(3)) != 0)) for l in wl.lights)
        for m in wl.models:
            batchers += bool(m.dynamic_batch_models)
            batched_static += m.auto_batch_static
            hidden += not m.visible
            blend_off += sum(p.material is not None and not p.material.blend_source_over for p in m.mesh.mesh_parts)
            skinned += m.skinned
            shape_keys += m.shape_key_weights is not None
            groups += m.light_group is not None
            for p in m.mesh.mesh_parts:
                if p.material is not None:
                    billboards += p.material.billboard
                    unbillboarded += p.material.depth_mode == A.DEPTH_MODE_UNBILLBOARDED
                    sort_modes.add(p.material.triangle_sort_mode)
                    transparency.add(p.material.transparency_mode)
                    lv_modes.add(p.material.light_volume_shading_mode)
    assert sort_modes == {A.SORT_BACK_TO_FRONT, A.SORT_FRONT_TO_BACK, A.SORT_NONE}
    assert len(transparency) == 4 and len(lv_modes) == 4
    assert skinned and shape_keys and groups and ortho and batched_static and blend_off and offset and hidden and normals
    assert billboards and unbillboarded and batchers
