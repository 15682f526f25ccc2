"""Seeded small scenes that turn the path's options on and off together, for
GPU-vs-oracle parity beyond the measured workloads: camera projection and
depth settings, snapping, perspective-corrected UVs, fog modes, every light
type (LightVolume shading modes included), light groups, sort modes,
backface culling, transparency modes and render order, lighting modes,
shadeless and hidden parts, skinning, shape keys, non-uniform scale and
frustum culling on and off."""
import math

import numpy as np

from tetra3d_amd import _abi as A
from tetra3d_amd import workloads as W
from tetra3d_amd.scene import (AmbientLight, Camera, DirectionalLight, LightVolume, Material, Mesh, MeshPart, Model,
                               PointLight, Scene, World, quat_to_matrix)

f32 = np.float32


def _rotation(rng):
    q = rng.normal(size=4)
    q /= np.linalg.norm(q)
    return quat_to_matrix(*q.astype(f32))


def _two_part_mesh(rng, base, name):
    """Split a mesh's triangles into two MeshParts with their own materials."""
    T = base.triangle_count
    cut = int(rng.integers(1, T)) if T > 1 else T
    parts = [MeshPart(base, Material(name + "_a"), 0, cut)]
    if cut < T:
        parts.append(MeshPart(base, Material(name + "_b"), cut, T - cut))
    base.mesh_parts = parts
    return base


def _mesh(rng, kind, name):
    if kind == "sphere":
        m = W.icosphere(int(rng.integers(1, 3)), float(rng.uniform(0.5, 2)), name=name,
                        color=None if rng.random() < 0.3 else (1, 1, 1, 1))
        if rng.random() < 0.5:   # per-vertex colours
            m.colors = np.ascontiguousarray(np.c_[rng.random((m.vertex_count, 3)), np.ones(m.vertex_count)], dtype=f32)
    elif kind == "grid":
        n = int(rng.integers(3, 12))
        m = W.grid_mesh(W.SplitMix64(int(rng.integers(1 << 30))), n, n, 0.5, 0.7, name=name)
    elif kind == "cube":
        m = W.cube_mesh(*rng.uniform(0.5, 2, 3))
    else:                     # one triangle: every key equal, the range==0 quirk of the sort
        m = Mesh(name, [[-1, -1, 0], [1, -1, 0], [0, 1, 0]], [[0, 0, 1]] * 3, [[0, 1, 2]],
                 uvs=[[0, 0], [1, 0], [0.5, 1]])
    if m.uvs is None and rng.random() < 0.5:
        m.uvs = np.ascontiguousarray(rng.random((m.vertex_count, 2)), dtype=f32)
    if rng.random() < 0.4 and m.triangle_count > 1:
        m = _two_part_mesh(rng, m, name)
    return m


def _material(rng, mat):
    mat.color = np.array([*rng.uniform(0.3, 1, 3), rng.choice([1.0, 1.0, 0.5, 0.0])], dtype=f32)
    mat.backface_culling = bool(rng.random() < 0.7)
    mat.triangle_sort_mode = int(rng.choice([A.SORT_BACK_TO_FRONT, A.SORT_FRONT_TO_BACK, A.SORT_NONE]))
    mat.shadeless = bool(rng.random() < 0.15)
    mat.fogless = bool(rng.random() < 0.2)
    mat.lighting_mode = int(rng.choice([A.LIGHTING_DEFAULT, A.LIGHTING_FIXED_NORMALS, A.LIGHTING_DOUBLE_SIDED]))
    mat.light_volume_shading_mode = int(rng.integers(0, 4))
    mat.transparency_mode = int(rng.integers(0, 4))
    mat.render_order = int(rng.integers(-1, 2))
    mat.visible = bool(rng.random() < 0.95)
    if rng.random() < 0.5:
        mat.texture_size = (float(rng.choice([16, 64, 256])), float(rng.choice([16, 128])))


def _skin(rng, model, bones=6):
    mesh = model.mesh
    V = mesh.vertex_count
    idx = np.stack([rng.permutation(bones)[:4] for _ in range(V)]).astype(np.uint16)
    w = rng.random((V, 4))
    w[rng.random((V, 4)) < 0.3] = 0
    w[:, 0] += 0.05
    w = (w / w.sum(axis=1, keepdims=True)).astype(f32)
    if rng.random() < 0.3:
        w[: V // 3] = np.array([1, 0, 0, 0], dtype=f32)   # single full-weight influence
    mesh.bone_indices, mesh.bone_weights = idx, w
    place = model.transform()
    pal = []
    for _ in range(bones):
        r = W.axis_angle_matrix(W.unit_vectors(W.SplitMix64(int(rng.integers(1 << 30))), 1)[0],
                                float(rng.uniform(0, 0.6)))
        pal.append(W.mat_mult(r, place))
    model.skinned = True
    model.bone_influences = np.stack(pal).astype(f32)


def _lights(rng, lo, hi):
    lights = []
    for k in range(int(rng.integers(1, 6))):
        t = rng.choice(["point", "point", "dir", "amb", "vol"])
        if t == "point":
            l = PointLight("p%d" % k, *rng.uniform(0.2, 1, 3), float(rng.uniform(0.5, 2)))
            l.set_local_position(*rng.uniform(lo - 2, hi + 2))
            l.range = f32(rng.choice([0.0, rng.uniform(2, 15)]))
        elif t == "dir":
            l = DirectionalLight("d%d" % k, *rng.uniform(0.2, 1, 3), float(rng.uniform(0.5, 1.5)))
            l.set_local_rotation(_rotation(rng))
        elif t == "amb":
            l = AmbientLight("a%d" % k, *rng.uniform(0.2, 1, 3), float(rng.uniform(0.1, 0.5)))
        else:
            l = LightVolume("v%d" % k, lo - 1, hi + 1, (2.0, 2.0, 2.0))
            l.cells[:, :3] = rng.random((len(l.cells), 3)).astype(f32)
            l.light_bias = rng.uniform(-0.3, 0.3, 3).astype(f32)
        l.visible = bool(rng.random() < 0.9)
        lights.append(l)
    return lights


def feature_scene(seed):
    rng = np.random.default_rng(seed)
    lo, hi = np.array([-6, -4, -14], dtype=f32), np.array([6, 4, -3], dtype=f32)
    cam = Camera("cam", int(rng.choice([320, 640, 1280])), int(rng.choice([240, 360, 720])))
    cam.perspective = bool(rng.random() < 0.8)
    cam.fov = f32(rng.uniform(40, 90))
    cam.near = f32(rng.choice([0.05, 0.1, 1.0]))
    cam.far = f32(rng.choice([30, 60, 300]))
    cam.ortho_scale = f32(rng.uniform(8, 30))
    cam.render_depth = bool(rng.random() < 0.7)
    cam.perspective_corrected_texture_mapping = bool(rng.random() < 0.5)
    cam.depth_margin = f32(rng.choice([0.0, 0.04, 0.2]))
    cam.vertex_snapping = f32(rng.choice([0, 0, 0, 0.25, 8]))
    if rng.random() < 0.3:
        cam.set_local_rotation(_rotation(rng) if rng.random() < 0.2 else W.axis_angle_matrix((0, 1, 0), 0.3))
    world = None
    if rng.random() < 0.9:
        world = World()
        world.lighting_on = bool(rng.random() < 0.9)
        world.fog_on = bool(rng.random() < 0.5)
        world.fog_mode = int(rng.integers(0, 4))
        world.fog_color = np.array([*rng.random(3), 1], dtype=f32)
        world.fog_range = np.array(sorted(rng.uniform(0, 1, 2)), dtype=f32)
        world.ambient_light.energy = f32(rng.uniform(0, 0.5))
        world.ambient_light.visible = bool(rng.random() < 0.9)
    lights = _lights(rng, lo, hi)
    models = []
    for i in range(int(rng.integers(4, 10))):
        kind = rng.choice(["sphere", "grid", "cube", "tri"])
        mesh = _mesh(rng, kind, "m%d" % i)
        if kind != "tri" and rng.random() < 0.3:        # two shape keys
            K = 2
            mesh.shape_key_positions = np.ascontiguousarray(rng.normal(0, 0.1, (K, mesh.vertex_count, 3)), dtype=f32)
            mesh.shape_key_normals = np.ascontiguousarray(rng.normal(0, 0.1, (K, mesh.vertex_count, 3)), dtype=f32)
        for p in mesh.mesh_parts:
            if p.material is None or rng.random() < 0.9:
                p.material = p.material or Material("mat%d" % i)
                _material(rng, p.material)
            else:
                p.material = None
        m = Model("model%d" % i, mesh)
        m.set_local_position(*rng.uniform(lo, hi))
        m.set_local_rotation(_rotation(rng))
        s = rng.uniform(0.5, 1.5, 3) if rng.random() < 0.3 else np.full(3, rng.uniform(0.5, 1.5))
        m.set_local_scale(*s)
        m.color = np.array([*rng.uniform(0.5, 1, 3), rng.choice([1.0, 0.7])], dtype=f32)
        m.shadeless = bool(rng.random() < 0.1)
        m.frustum_culling = bool(rng.random() < 0.8)
        if mesh.shape_key_positions is not None:
            m.shape_key_weights = rng.uniform(0, 1, len(mesh.shape_key_positions)).astype(f32)
        if kind in ("sphere", "grid") and mesh.shape_key_positions is None and rng.random() < 0.35:
            _skin(rng, m)
        if rng.random() < 0.2 and lights:
            m.light_group = [l for l in lights if not isinstance(l, LightVolume) and rng.random() < 0.6]
        models.append(m)
    # further options, drawn from a stream of their own so the draws above stay as they were
    extra = np.random.default_rng(10_000 + seed)
    for m in models:
        m.auto_batch_static = bool(extra.random() < 0.05)           # skipped (camera.go:2067-2071)
        m.visible = bool(extra.random() < 0.95)
        for p in m.mesh.mesh_parts:
            if p.material is not None and extra.random() < 0.15:
                p.material.blend_source_over = False                # a transparent pair (camera.go:1420-1428)
    for l in lights:
        if isinstance(l, LightVolume) and extra.random() < 0.5:
            l.offset = extra.uniform(-1.5, 1.5, 3).astype(f32)
    cam.render_normals = bool(extra.random() < 0.35)                # the normal vertex list (camera.go:1710-1722)
    for m in models:                                                 # billboards (model.go:634-706)
        for p in m.mesh.mesh_parts:
            if p.material is not None and extra.random() < 0.25:
                p.material.billboard = True
                p.material.billboard_up = int(extra.integers(0, 2))
                p.material.billboard_lock = tuple(bool(x) for x in extra.random(3) < 0.3)
                if not m.skinned:
                    p.material.depth_mode = int(extra.integers(0, 2))
    if extra.random() < 0.5 and len(models) > 3:                     # dynamic batches (model.go:217-247)
        owner = models[int(extra.integers(0, len(models)))]
        others = [m for m in models if m is not owner]
        for i in extra.choice(len(others), size=int(extra.integers(1, min(len(others), 6) + 1)), replace=False):
            owner.dynamic_batch_add(int(extra.integers(0, len(owner.mesh.mesh_parts))), others[int(i)])
    cells = np.random.default_rng(20_000 + seed)                     # volume cell lights (light.go:650-653)
    for l in lights:
        if isinstance(l, LightVolume) and cells.random() < 0.6:
            add_cell_lights(cells, l, lo, hi)
    return W.Workload("features_%d" % seed, cam, Scene("features", world), lights, models)


def add_cell_lights(rng, vol, lo, hi, n_lights=None, frac=0.3):
    """Point, directional and ambient lights (some hidden, which cell lights
    ignore) named by about `frac` of the volume's cells, one to three each,
    repeats allowed (LightVolumeCell.SetUseLights, light.go:1508-1513)."""
    nested = []
    for k in range(n_lights or int(rng.integers(1, 5))):
        t = rng.choice(["point", "point", "dir", "amb"])
        if t == "point":
            l = PointLight("%s.p%d" % (vol.name, k), *rng.uniform(0.2, 1, 3), float(rng.uniform(0.5, 2)))
            l.set_local_position(*rng.uniform(lo - 2, hi + 2))
            l.range = f32(rng.choice([0.0, rng.uniform(1, 6)]))
        elif t == "dir":
            l = DirectionalLight("%s.d%d" % (vol.name, k), *rng.uniform(0.2, 1, 3), float(rng.uniform(0.5, 1.5)))
            l.set_local_rotation(_rotation(rng))
        else:
            l = AmbientLight("%s.a%d" % (vol.name, k), *rng.uniform(0.2, 1, 3), float(rng.uniform(0.1, 0.5)))
        l.visible = bool(rng.random() < 0.7)
        nested.append(l)
    for c in np.flatnonzero(rng.random(len(vol.cells)) < frac):
        vol.set_use_lights(int(c), *[nested[int(i)] for i in rng.integers(0, len(nested), int(rng.integers(1, 4)))])
    return nested


def cell_lights_scene(seed, many=False):
    """Models inside a LightVolume whose cells name lights, in every shading
    mode, rigid and skinned, the volume in the scene lights and in a
    LightGroup; `many` gives the volume 80 cell lights, so a pair's light list
    passes what the vertex pass stages in LDS."""
    rng = np.random.default_rng(seed)
    lo, hi = np.array([-6, -4, -14], dtype=f32), np.array([6, 4, -3], dtype=f32)
    cam = Camera("cam", 640, 360)
    cam.far = f32(60)
    world = World()
    world.ambient_light.energy = f32(0.2)
    vol = LightVolume("vol", lo - 1, hi + 1, (2.0, 2.0, 2.0))
    vol.cells[:, :3] = rng.random((len(vol.cells), 3)).astype(f32)
    add_cell_lights(rng, vol, lo, hi, n_lights=80 if many else 6, frac=0.8)
    sun = DirectionalLight("sun", 0.9, 0.8, 0.7, 1.0)
    sun.set_local_rotation(_rotation(rng))
    lights = [sun, vol]
    models = []
    for i in range(8):
        kind = ["sphere", "grid", "cube", "tri"][i % 4]
        mesh = _mesh(rng, kind, "m%d" % i)
        for p in mesh.mesh_parts:
            p.material = p.material or Material("mat%d" % i)
            _material(rng, p.material)
            p.material.light_volume_shading_mode = [A.LV_PER_VERTEX_WITHOUT_NORMAL, A.LV_PER_VERTEX_WITH_NORMAL,
                                                    A.LV_PER_OBJECT, A.LV_OFF][i % 4 if i < 4 else int(rng.integers(0, 4))]
            p.material.blend_source_over = True
        m = Model("model%d" % i, mesh)
        m.set_local_position(*rng.uniform(lo, hi))
        m.set_local_rotation(_rotation(rng))
        m.frustum_culling = False
        if kind in ("sphere", "grid") and i >= 4:
            _skin(rng, m)
        if i % 3 == 2:
            m.light_group = [vol, sun] if i % 2 else [vol]
        models.append(m)
    return W.Workload("cell_lights_%d" % seed, cam, Scene("cells", world), lights, models)


def pair_order_scene(seed, n_models=24):
    """Models stacked on a few shared positions (equal camera distances), one
    render order for the solids and several for the transparent parts: the
    order of render pairs then rests on the stable-sort ties."""
    rng = np.random.default_rng(seed)
    cam = Camera("cam", 640, 360)
    cam.render_depth = bool(seed % 2)
    spots = rng.uniform([-4, -3, -12], [4, 3, -5], (4, 3)).astype(f32)
    lights = [DirectionalLight("sun", 1, 1, 1, 1.0)]
    lights[0].set_local_rotation(W.axis_angle_matrix((1, 0, 0), -0.6))
    models = []
    for i in range(n_models):
        mesh = W.icosphere(1, float(rng.uniform(0.5, 1.5)), name="s%d" % i)
        mat = Material("mat%d" % i)
        transparent = rng.random() < 0.4
        mat.color = np.array([*rng.uniform(0.3, 1, 3), 0.5 if transparent else 1.0], dtype=f32)
        mat.render_order = int(rng.integers(-1, 2)) if transparent else 0
        mat.backface_culling = True
        for p in mesh.mesh_parts:
            p.material = mat
        m = Model("model%d" % i, mesh)
        m.set_local_position(*spots[rng.integers(0, len(spots))])
        m.set_local_rotation(_rotation(rng))
        models.append(m)
    world = World()
    world.ambient_light.energy = f32(0.3)
    return W.Workload("pair_order_%d" % seed, cam, Scene("pair_order", world), lights, models)


def many_bones_scene(seed):
    """Skinned models whose palettes straddle the LDS staging limit of the
    vertex pass (64 bones): 63, 64, 65 and 96 bones, so both the staged and
    the global-memory palette reads are exercised."""
    rng = np.random.default_rng(seed)
    cam = Camera("cam", 640, 360)
    lights = [DirectionalLight("sun", 1, 1, 1, 1.0)]
    lights[0].set_local_rotation(W.axis_angle_matrix((1, 0, 0), -0.6))
    models = []
    for i, bones in enumerate([63, 64, 65, 96]):
        mesh = W.icosphere(2, 1.0, name="b%d" % i)
        m = Model("skinned%d" % i, mesh)
        m.set_local_position(-4.5 + 3 * i, 0, -8)
        m.set_local_rotation(_rotation(rng))
        _skin(rng, m, bones=bones)
        models.append(m)
    world = World()
    world.ambient_light.energy = f32(0.2)
    return W.Workload("many_bones_%d" % seed, cam, Scene("many_bones", world), lights, models)


def random_poses(rng, n, tree=True):
    """An armature of `n` bones as poses, parents first: small local offsets,
    rotations and near-unit scales (one mirrored), parents drawn among the
    earlier bones (tree) or each bone under the previous one (chain), with
    inverse bind matrices of a rest pose."""
    poses = (A.BonePose * max(1, n))()
    for b in range(n):
        p = poses[b]
        p.position[:] = [float(x) for x in rng.uniform(-0.08, 0.08, 3).astype(f32)]
        s = rng.uniform(0.9, 1.1, 3)
        if b == 2:
            s[1] = -s[1]
        p.scale[:] = [float(x) for x in s.astype(f32)]
        p.parent = -1 if b == 0 else (int(rng.integers(-1, b)) if tree else b - 1)
        axis = W.unit_vectors(W.SplitMix64(int(rng.integers(1 << 30))), 1)[0]
        p.rotation = A.mat4(W.axis_angle_matrix(axis, float(rng.uniform(0, 0.4))))
        rest = W.axis_angle_matrix(axis, float(rng.uniform(0, 0.2)))
        rest = np.array(rest, dtype=np.float64)
        rest[3, :3] = rng.uniform(-0.05, 0.05, 3)
        p.inverse_bind = A.mat4(np.linalg.inv(rest).astype(f32))
    return poses


def huge_armature_scene(seed, bones=5000):
    """One skinned model whose armature has more bones than the palette pass
    walks level by level (4096), so its parent chain is walked bone by bone,
    beside a 12-bone one."""
    rng = np.random.default_rng(seed)
    cam = Camera("cam", 640, 360)
    lights = [DirectionalLight("sun", 1, 1, 1, 1.0)]
    models = []
    for i, n in enumerate([bones, 12]):
        m = Model("huge%d" % i, W.icosphere(2, 1.0, name="huge%d" % i))
        m.set_local_position(-1.5 + 3 * i, 0, -6)
        _skin(rng, m, bones=n)
        m.bone_poses = random_poses(rng, n, True)
        m.armature_parent = np.array(m.transform(), dtype=f32)
        m.bone_influences = None
        models.append(m)
    world = World()
    world.ambient_light.energy = f32(0.2)
    return W.Workload("huge_armature_%d" % seed, cam, Scene("huge", world), lights, models)


def posed_scene(seed):
    """Skinned models whose palettes are computed from bone poses: a tree of
    12 bones on a two-part mesh (one palette job shared by two render pairs),
    a 300-bone tree (past the 256 bones the palette pass keeps in LDS), a
    40-bone chain without an armature parent, one bone, and a model with a
    host palette beside them."""
    rng = np.random.default_rng(seed)
    cam = Camera("cam", 640, 360)
    lights = [DirectionalLight("sun", 1, 1, 1, 1.0), PointLight("lamp", 1, 0.8, 0.6, 2.0)]
    lights[0].set_local_rotation(W.axis_angle_matrix((1, 0, 0), -0.6))
    lights[1].set_local_position(0, 3, -5)
    models = []
    specs = [("two_part", 12, True, True), ("big_tree", 300, True, True), ("chain", 40, False, False),
             ("one_bone", 1, True, True), ("host_palette", 6, None, None)]
    for i, (kind, bones, tree, rooted) in enumerate(specs):
        if kind == "two_part":
            mesh = _two_part_mesh(rng, W.icosphere(2, 1.0, name="pose_base"), "pose%d" % i)
        else:
            mesh = W.icosphere(2, 1.0, name="pose%d" % i)
        m = Model("posed_%s" % kind, mesh)
        m.set_local_position(-6 + 3 * i, 0, -9)
        m.set_local_rotation(_rotation(rng))
        _skin(rng, m, bones=max(4, bones))
        if bones < 4:                                   # every vertex on bone 0
            V = mesh.vertex_count
            mesh.bone_indices = np.zeros((V, 4), dtype=np.uint16)
            mesh.bone_weights = np.tile(np.array([1, 0, 0, 0], dtype=f32), (V, 1))
        if tree is not None:
            m.bone_poses = random_poses(rng, bones, tree)
            if rooted:
                m.armature_parent = np.array(m.transform(), dtype=f32)
            else:
                m.bone_poses[0].position[:] = [float(x) for x in np.asarray(m.transform(), dtype=f32)[3, :3]]
            m.bone_influences = None
        models.append(m)
    world = World()
    world.ambient_light.energy = f32(0.2)
    return W.Workload("posed_%d" % seed, cam, Scene("posed", world), lights, models)
