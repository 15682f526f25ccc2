"""The raster sink on the GPU against the oracle's rasteriser, image for
image and bit for bit: the closed-form scenes of the oracle raster tests, the
seeded feature scenes (textures, both filters, AlphaClip, fog curves and
dithering, transparency, dynamic batches) and the mixed-512 frame at
1920 x 1080, plus triangles far off screen and whole-screen ones."""
import numpy as np
import pytest

from tetra3d_amd import _abi as A
from tetra3d_amd import workloads as W
from tetra3d_amd.scene import Scene, Texture, with_batched
from tests import test_oracle_raster as R
from tests.feature_scenes import feature_scene

pytestmark = pytest.mark.gpu


def _both(gpu_ctx, oracle, cam, scene, lights, models):
    for m in {id(m.mesh): m.mesh for m in with_batched(models) if m.mesh is not None}.values():
        gpu_ctx.upload_mesh(m)
        if id(m) not in oracle._handles:
            oracle.register_mesh(m, m._handles["gpu"])
    gpu_ctx.upload_textures(models)
    gpu_ctx.render(cam, scene, lights, models, download=False)
    gpu_ctx.rasterize()
    gc, gd = gpu_ctx.download_images()
    _, oc, od = oracle.rasterize(cam, scene, lights, models)
    return gc, gd, oc, od


def _same(gc, gd, oc, od, what):
    bad = np.argwhere(np.any(gc != oc, axis=-1))
    assert len(bad) == 0, "%s: %d colour pixels differ, first %s gpu %s oracle %s" % (
        what, len(bad), bad[:3].tolist(), [gc[tuple(b)].tolist() for b in bad[:3]], [oc[tuple(b)].tolist() for b in bad[:3]])
    assert (gd is None) == (od is None)
    if gd is not None:
        bad = np.argwhere(np.any(gd != od, axis=-1))
        assert len(bad) == 0, "%s: %d depth pixels differ, first %s" % (what, len(bad), bad[:3].tolist())


def _scenes():
    """The oracle raster tests' scenes, rebuilt."""
    out = []
    out.append(("rect", None, [R._flat(R._square(100, 40, 164, 88), color=(1, 1, 1, 0.5))]))
    cx, cy = R._px(201.3, 97.7)
    ring = [R._px(201.3 + 61.1 * np.cos(k * np.pi / 4 + 0.3), 97.7 + 47.9 * np.sin(k * np.pi / 4 + 0.3)) for k in range(8)]
    out.append(("fan", None, [R._flat([((cx, cy), ring[k], ring[(k + 1) % 8]) for k in range(8)], color=(1, 1, 1, 0.5))]))
    out.append(("depth_order", None, [R._flat(R._square(100, 40, 164, 88), z=-3, color=(1, 0, 0, 1)),
                                      R._flat(R._square(120, 50, 184, 98), z=-6, color=(0, 1, 0, 1))]))
    tri = R._square(100, 40, 164, 88)[0]
    both = R._flat([tri, tri], z=-4, color=(0, 0, 1, 1))
    both.mesh.positions[3:, 2] = -2.0
    out.append(("within_flush", None, [both]))
    out.append(("transparent", None, [R._flat(R._square(100, 40, 164, 88), z=-3, color=(1, 0, 0, 0.5)),
                                      R._flat(R._square(100, 40, 164, 88), z=-6, color=(0, 1, 0, 1))]))
    w = R._world(fog_on=True, fog_mode=A.FOG_ADD, fog_color=np.array([0.25, 0.5, 1, 1], dtype=np.float32),
                 fog_range=np.array([0.0, 0.5], dtype=np.float32), dithered_fog_size=1.0)
    out.append(("dither", w, [R._flat(R._square(96, 40, 160, 88), color=(0.5, 0.5, 0.5, 1))]))
    uvs = np.array([(0, 1), (1, 1), (1, 0), (0, 1), (1, 0), (0, 0)], dtype=np.float32)
    m = R._flat(R._square(100, 40, 164, 104), texture=R._checker(), uvs=uvs)
    m.mesh.mesh_parts[0].material.transparency_mode = A.TRANSPARENCY_ALPHA_CLIP
    out.append(("clip", None, [m]))
    # far off-screen corners and a whole-screen pair
    big = R._flat([((-40, -30), (55, -20), (3, 70)), ((-3, -3), (3, -3), (0, 3))], z=-5, color=(0.3, 0.6, 0.9, 1))
    out.append(("huge", None, [big]))
    return out


def test_closed_form_scenes(gpu_ctx, oracle):
    for name, w, models in _scenes():
        print("raster scene", name, flush=True)
        cam = R.ortho_camera()
        _same(*_both(gpu_ctx, oracle, cam, Scene("s", w if w is not None else R._world()), [], models), name)


def _dress(wl, seed):
    """Raster options on a feature scene: SourceOver everywhere (others are
    refused), random textures for the textured materials in both filters and
    AlphaClip on some, a fog curve, dithering and a clear colour."""
    rng = np.random.default_rng(50_000 + seed)
    for m in with_batched(wl.models):
        for p in m.mesh.mesh_parts:
            mat = p.material
            if mat is None:
                continue
            mat.blend_source_over = True
            if mat.texture_size is not None and rng.random() < 0.8:
                tw, th = int(mat.texture_size[0]), int(mat.texture_size[1])
                img = rng.integers(0, 256, (th, tw, 4), dtype=np.uint8)
                img[..., :3] = (img[..., :3].astype(np.uint16) * img[..., 3:4] // 255).astype(np.uint8)
                mat.texture = Texture(img)
                mat.texture_filter = int(rng.integers(0, 2))
                if rng.random() < 0.3:
                    mat.transparency_mode = A.TRANSPARENCY_ALPHA_CLIP
            mat.fogless = bool(rng.random() < 0.2)
    w = wl.scene.world
    if w is not None:
        w.fog_curve = int(rng.integers(0, 3))
        w.dithered_fog_size = float(rng.choice([0.0, 0.0, 1.0, 2.5]))
        w.clear_color = rng.uniform(0, 1, 4).astype(np.float32)
    return wl


def test_feature_scenes(gpu_ctx, oracle):
    drawn = 0
    for seed in range(24):
        wl = _dress(feature_scene(seed), seed)
        print("raster feature scene", seed, flush=True)
        gpu_ctx.prepare(wl)
        oracle.render_workload(wl)   # registers the meshes, volumes and animations under the device's handles
        gc, gd, oc, od = _both(gpu_ctx, oracle, wl.camera, wl.scene, wl.lights, wl.models)
        _same(gc, gd, oc, od, "feature scene %d" % seed)
        drawn += int(np.count_nonzero(np.any(gc != gc[0, 0], axis=-1)))
    assert drawn > 0


def test_mixed_512_full_size(gpu_ctx, oracle):
    wl = W.mixed_scene()
    print("raster mixed_512, %d triangles" % wl.triangle_count, flush=True)
    gpu_ctx.prepare(wl)
    oracle.render_workload(wl)
    gc, gd, oc, od = _both(gpu_ctx, oracle, wl.camera, wl.scene, wl.lights, wl.models)
    _same(gc, gd, oc, od, "mixed_512")
    assert np.count_nonzero(gd[..., 3]) > 100000
