"""glTF meshes put in the resident pool by t3d_gltf_upload render bit for bit
as the oracle renders the same decoded arrays: the reference's example
exports (vertex colours of 16 bits and float, several primitives per mesh, a
skinned mesh) and a synthetic mesh with 8-bit colours and morph targets."""
import os

import numpy as np
import pytest

from tests.gltf_writer import Builder
from tests.parity import compare_frames
from tetra3d_amd import workloads as W
from tetra3d_amd.gltf import Document
from tetra3d_amd.scene import AmbientLight, Camera, DirectionalLight, Model, PointLight, Scene, World

f32 = np.float32
FIXTURES = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "gltf")


def _scene(doc, rng):
    """Every mesh of the document as a model, side by side in front of the
    camera, scaled to a common size; skinned meshes get random palettes."""
    cam = Camera("cam", 640, 360)
    cam.set_local_position(0, 0, 12)
    lights = [DirectionalLight("sun", 1, 0.9, 0.8, 1.0), PointLight("p", 0.5, 0.7, 1.0, 2.0)]
    lights[0].set_local_rotation(W.axis_angle_matrix((1, 0, 0), -0.7))
    lights[1].set_local_position(2, 3, 4)
    lights[1].range = f32(20)
    models = []
    n = len(doc.meshes)
    for i, (mesh, info) in enumerate(zip(doc.meshes, doc.infos)):
        m = Model("m%d" % i, mesh)
        for part in mesh.mesh_parts:   # both sides drawn: flat exports seen from behind still draw
            part.material.backface_culling = False
        span = float(np.max(mesh.dimensions[1] - mesh.dimensions[0])) or 1.0
        s = 3.0 / span
        m.set_local_scale(s, s, s)
        m.set_local_position(-5 + 10 * (i + 0.5) / n, float(rng.uniform(-1, 1)), float(rng.uniform(-2, 0)))
        m.set_local_rotation(W.axis_angle_matrix((0.3, 1, 0.2), float(rng.uniform(0, 6.28))))
        if info.skinned:
            nb = int(mesh.bone_indices.max()) + 1
            m.skinned = True
            m.bone_influences = np.stack([W.axis_angle_matrix((0, 0, 1), float(a)) for a in rng.uniform(-0.3, 0.3, nb)]).astype(f32)
        if mesh.shape_key_positions is not None:
            m.shape_key_weights = rng.uniform(0, 1, len(mesh.shape_key_positions)).astype(f32)
        models.append(m)
    world = World()
    world.ambient_light.energy = f32(0.3)
    return W.Workload("gltf", cam, Scene("gltf", world), lights, models)


def _synthetic():
    b = Builder()
    rng = np.random.default_rng(3)
    ico = W.icosphere(2, 1.0)
    V = ico.vertex_count
    col = rng.integers(0, 256, (V, 4))
    d = rng.standard_normal((V, 3)).astype(f32) * f32(0.2)
    b.mesh("ico", [{"attributes": {"POSITION": b.accessor(ico.positions, "f32"), "NORMAL": b.accessor(ico.normals, "f32"),
                                   "COLOR_0": b.accessor(col, "u8", normalized=True)},
                    "indices": b.accessor(ico.indices.reshape(-1), "u16"), "targets": [{"POSITION": b.accessor(d, "f32")}]}],
           extras={"t3dVertexColorNames__": ["Col"]})
    return b.glb()


@pytest.mark.gpu
@pytest.mark.parametrize("name", sorted(f for f in os.listdir(FIXTURES) if f.endswith(".glb")) + ["synthetic"])
def test_gltf_meshes_render_as_the_oracle(gpu_ctx, oracle, name):
    data = _synthetic() if name == "synthetic" else open(os.path.join(FIXTURES, name), "rb").read()
    doc = Document(data)
    handles = doc.upload(gpu_ctx)
    print("gltf", name, "meshes", len(handles), flush=True)
    assert all(h for h in handles)
    wl = _scene(doc, np.random.default_rng(len(name)))
    g = gpu_ctx.render_workload(wl)
    s = compare_frames(g, oracle.render_workload(wl))
    assert s["survivors"] > 0
    for m in doc.meshes:
        gpu_ctx.free_mesh(m)
