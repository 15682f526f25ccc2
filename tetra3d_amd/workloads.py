"""Seeded synthetic scenes for the geometry path.

One PRNG (SplitMix64, seed 0x7e7a3d unless given) drives every random choice;
floats come from the top 24 bits of each draw.  The same arrays feed the GPU
path and the oracle, so both see byte-identical inputs.

Scenes (sizes as in the measurement plan):
  icosphere_scene  — 642-vertex shared-edge icosphere, 640x360, 3 lights
  reference_icosphere_scene — the same on NewIcosphereMesh(2) (1,272 vertices)
  cubes_scene      — 4,096 unit cubes in a 200^3 box, 1920x1080, 10 lights
  heightfield_scene— 1024x1024 heightfield, camera inside the terrain, 18 lights
  mixed_scene      — 512 models, V log-uniform in [24, 65536], 128 skinned
  crowd_scene      — 256 characters on one skinned mesh, 64-bone armatures as poses, animated
  serving_pool     — 10,000 models of the mixed distribution + a Poisson job trace
"""
import math

import numpy as np

from . import _abi as A
from .scene import (AmbientLight, Camera, DirectionalLight, LightVolume, Material, Mesh, MeshPart, Model,
                    PointLight, Scene, World, identity, mat_mult, mat_scale, mat_translate, quat_to_matrix,
                    rotation_from_forward, with_batched)

f32 = np.float32
DEFAULT_SEED = 0x7E7A3D
_GAMMA = np.uint64(0x9E3779B97F4A7C15)
_M1 = np.uint64(0xBF58476D1CE4E5B9)
_M2 = np.uint64(0x94D049BB133111EB)


class SplitMix64:
    """SplitMix64: output k (1-based) mixes seed + k*0x9E3779B97F4A7C15."""

    def __init__(self, seed=DEFAULT_SEED):
        self.seed = np.uint64(seed & 0xFFFFFFFFFFFFFFFF)
        self.counter = 0

    def u64(self, n):
        k = np.arange(self.counter + 1, self.counter + n + 1, dtype=np.uint64)
        self.counter += n
        with np.errstate(over="ignore"):
            z = self.seed + k * _GAMMA
            z = (z ^ (z >> np.uint64(30))) * _M1
            z = (z ^ (z >> np.uint64(27))) * _M2
            z = z ^ (z >> np.uint64(31))
        return z

    def floats(self, n):
        """n floats in [0, 1) from the top 24 bits."""
        return ((self.u64(n) >> np.uint64(40)).astype(np.float64) / float(1 << 24)).astype(f32)

    def uniform(self, lo, hi, n):
        return (np.float64(lo) + (np.float64(hi) - np.float64(lo)) * self.floats(n).astype(np.float64)).astype(f32)

    def ints(self, lo, hi, n):
        """n integers in [lo, hi)."""
        return (lo + (self.u64(n) % np.uint64(hi - lo)).astype(np.int64)).astype(np.int64)

    def child(self, tag):
        """An independent stream derived from this one's seed and a tag."""
        return SplitMix64(int(self.seed) ^ ((tag * 0xD1B54A32D192ED03) & 0xFFFFFFFFFFFFFFFF))


def uniform_quaternions(rng, n):
    """Shoemake's uniform random unit quaternions, as (x, y, z, w) float32."""
    u = rng.floats(3 * n).astype(np.float64).reshape(n, 3)
    a, b = np.sqrt(1 - u[:, 0]), np.sqrt(u[:, 0])
    t1, t2 = 2 * np.pi * u[:, 1], 2 * np.pi * u[:, 2]
    return np.stack([a * np.sin(t1), a * np.cos(t1), b * np.sin(t2), b * np.cos(t2)], axis=1).astype(f32)


def unit_vectors(rng, n):
    u = rng.floats(2 * n).astype(np.float64).reshape(n, 2)
    z = 2 * u[:, 0] - 1
    phi = 2 * np.pi * u[:, 1]
    r = np.sqrt(np.maximum(0, 1 - z * z))
    return np.stack([r * np.cos(phi), r * np.sin(phi), z], axis=1).astype(f32)


def axis_angle_matrix(axis, angle):
    """Rotation Matrix4 about a unit axis (the NewMatrix4Rotate form, matrix.go:101-128)."""
    x, y, z = (float(a) for a in axis)
    s, c = math.sin(angle), math.cos(angle)
    m = 1 - c
    r = np.eye(4, dtype=np.float64)
    r[0, 0] = m * x * x + c
    r[0, 1] = m * x * y + z * s
    r[0, 2] = m * z * x - y * s
    r[1, 0] = m * x * y - z * s
    r[1, 1] = m * y * y + c
    r[1, 2] = m * y * z + x * s
    r[2, 0] = m * z * x + y * s
    r[2, 1] = m * y * z - x * s
    r[2, 2] = m * z * z + c
    return r.astype(f32)


# ---------------------------------------------------------------------------
# meshes
# ---------------------------------------------------------------------------

_ICO_CACHE = {}


def _icosphere_topology(level):
    """Unit-sphere vertices (float64) and outward-wound faces of a level."""
    if level not in _ICO_CACHE:
        _ICO_CACHE[level] = _build_icosphere_topology(level)
    return _ICO_CACHE[level]


def icosphere(level, radius=1.0, name="icosphere", color=None):
    """Shared-edge icosphere: 10*4^level + 2 vertices, 20*4^level triangles,
    smooth normals, outward counter-clockwise winding (front faces survive the
    screen-space backface test)."""
    p, idx = _icosphere_topology(level)
    pos = (p * radius).astype(f32)
    nrm = p.astype(f32)
    cols = None
    if color is not None:
        cols = np.tile(np.asarray(color, dtype=f32), (len(pos), 1))
    uv = np.stack([0.5 + np.arctan2(p[:, 2], p[:, 0]) / (2 * np.pi), 0.5 + np.arcsin(np.clip(p[:, 1], -1, 1)) / np.pi],
                  axis=1).astype(f32)
    return Mesh(name, pos, nrm, idx.astype(np.uint32), colors=cols, uvs=uv)


def reference_icosphere_mesh(detail_level=2, name="Icosphere", color=(1, 1, 1, 1)):
    """NewIcosphereMesh(detail_level) as the reference builds it (mesh.go:1394-1525),
    in float32: the 12 golden-ratio vertices, detail_level + 1 subdivisions
    whose midpoint cache is keyed by the ordered index pair (so each edge's
    midpoint is made once per direction: 1,272 vertices at level 2),
    Unit() on every vertex, U/V set by index modulo 3, and AutoNormal
    (mesh.go:755-763): each triangle's normal (mesh.go:2030-2037, 2171-2179)
    assigned to its three vertices, the last triangle winning."""
    def unit(x, y, z):                                  # Vector3.Unit, vector.go:258-266
        l = f32(math.sqrt(float(f32(f32(x * x + y * y) + z * z))))
        if l < f32(1e-8) or l == f32(1):
            return x, y, z
        return f32(x / l), f32(y / l), f32(z / l)

    t = f32(f32(f32(1) + f32(math.sqrt(5.0))) / f32(2))
    one, zero = f32(1), f32(0)
    verts = [(-one, t, zero), (one, t, zero), (-one, -t, zero), (one, -t, zero),
             (zero, -one, t), (zero, one, t), (zero, -one, -t), (zero, one, -t),
             (t, zero, -one), (t, zero, one), (-t, zero, -one), (-t, zero, one)]
    idx = [0, 11, 5, 0, 5, 1, 0, 1, 7, 0, 7, 10, 0, 10, 11, 1, 5, 9, 5, 11, 4, 11, 10, 2, 10, 7, 6, 7, 1, 8,
           3, 9, 4, 3, 4, 2, 3, 2, 6, 3, 6, 8, 3, 8, 9, 4, 9, 5, 2, 4, 11, 6, 2, 10, 8, 6, 7, 9, 8, 1]
    cache = {}

    def middle(i0, i1):
        if (i0, i1) in cache:
            return cache[(i0, i1)]
        a, b = verts[i0], verts[i1]
        verts.append(tuple(f32(f32(a[k] + b[k]) / f32(2)) for k in range(3)))   # Add, then Divide(2)
        cache[(i0, i1)] = len(verts) - 1
        return len(verts) - 1

    if detail_level > 0:
        for _ in range(detail_level + 1):
            new = []
            for i in range(0, len(idx), 3):
                ab, bc, ca = middle(idx[i], idx[i + 1]), middle(idx[i + 1], idx[i + 2]), middle(idx[i + 2], idx[i])
                new += [idx[i], ab, ca, idx[i + 1], bc, ab, idx[i + 2], ca, bc, ab, bc, ca]
            idx = new
    pos = np.array([unit(*v) for v in verts], dtype=f32)
    uv = np.zeros((len(pos), 2), dtype=f32)
    uv[1::3, 0] = 1
    uv[2::3, 1] = 1
    nrm = np.zeros_like(pos)
    tris = np.array(idx, dtype=np.uint32).reshape(-1, 3)
    for a, b, c in tris:
        p1, p2, p3 = pos[a], pos[b], pos[c]
        v0 = [f32(p2[k] - p1[k]) for k in range(3)]
        v1 = [f32(p3[k] - p2[k]) for k in range(3)]
        cz = f32(f32(v0[0] * v1[1]) - f32(v1[0] * v0[1]))         # Vector3.Cross, vector.go:164-171
        cy = f32(f32(v0[2] * v1[0]) - f32(v1[2] * v0[0]))
        cx = f32(f32(v0[1] * v1[2]) - f32(v1[1] * v0[2]))
        n = unit(cx, cy, cz)
        if n == (0, 0, 0):
            n = tuple(nrm[a])
        nrm[a] = nrm[b] = nrm[c] = n
    cols = np.tile(np.asarray(color, dtype=f32), (len(pos), 1)) if color is not None else None
    return Mesh(name, pos, nrm, tris, colors=cols, uvs=uv)


def reference_icosphere_scene(seed=DEFAULT_SEED, detail_level=2):
    """The icosphere configuration on the reference's own generator,
    NewIcosphereMesh(2): 1,272 vertices, 1,280 triangles, flat normals."""
    wl = icosphere_scene(seed)
    model = wl.models[0]
    ref = Model("icosphere_ref", reference_icosphere_mesh(detail_level))
    ref.set_local_rotation(model.rotation)
    return Workload("icosphere_ref_1272v", wl.camera, wl.scene, wl.lights, [ref])


def _build_icosphere_topology(level):
    t = (1 + math.sqrt(5)) / 2
    verts = [(-1, t, 0), (1, t, 0), (-1, -t, 0), (1, -t, 0), (0, -1, t), (0, 1, t), (0, -1, -t), (0, 1, -t),
             (t, 0, -1), (t, 0, 1), (-t, 0, -1), (-t, 0, 1)]
    verts = [np.array(v, dtype=np.float64) / np.linalg.norm(v) for v in verts]
    faces = [(0, 11, 5), (0, 5, 1), (0, 1, 7), (0, 7, 10), (0, 10, 11), (1, 5, 9), (5, 11, 4), (11, 10, 2),
             (10, 7, 6), (7, 1, 8), (3, 9, 4), (3, 4, 2), (3, 2, 6), (3, 6, 8), (3, 8, 9), (4, 9, 5), (2, 4, 11),
             (6, 2, 10), (8, 6, 7), (9, 8, 1)]
    for _ in range(level):
        cache = {}

        def mid(a, b):
            key = (min(a, b), max(a, b))
            if key not in cache:
                m = verts[a] + verts[b]
                verts.append(m / np.linalg.norm(m))
                cache[key] = len(verts) - 1
            return cache[key]

        nf = []
        for a, b, c in faces:
            ab, bc, ca = mid(a, b), mid(b, c), mid(c, a)
            nf += [(a, ab, ca), (b, bc, ab), (c, ca, bc), (ab, bc, ca)]
        faces = nf
    p = np.array(verts)
    idx = np.array(faces, dtype=np.int64)
    n = np.cross(p[idx[:, 1]] - p[idx[:, 0]], p[idx[:, 2]] - p[idx[:, 0]])
    flip = np.einsum("ij,ij->i", n, p[idx].sum(axis=1)) < 0
    idx[flip] = idx[flip][:, [0, 2, 1]]
    p.setflags(write=False)
    idx.setflags(write=False)
    return p, idx


def cube_mesh(w=1.0, h=1.0, d=1.0):
    """The 24-vertex / 12-triangle cube of NewCubeMesh (mesh.go:1251-1342),
    with AutoNormal's last-writer face normals (mesh.go:755-764)."""
    w, h, d = f32(w) / f32(2), f32(h) / f32(2), f32(d) / f32(2)
    v = [(-w, h, -d, 0, 0), (w, h, d, 1, 1), (w, h, -d, 1, 0), (-w, h, d, 0, 1),
         (w, -h, -d, 1, 0), (w, -h, d, 1, 1), (-w, -h, -d, 0, 0), (-w, -h, d, 0, 1),
         (-w, h, d, 0, 0), (w, -h, d, 1, 1), (w, h, d, 1, 0), (-w, -h, d, 0, 1),
         (w, h, -d, 1, 0), (w, -h, -d, 1, 1), (-w, h, -d, 0, 0), (-w, -h, -d, 0, 1),
         (w, h, -d, 1, 0), (w, h, d, 1, 1), (w, -h, -d, 0, 0), (w, -h, d, 0, 1),
         (-w, -h, -d, 0, 0), (-w, h, d, 1, 1), (-w, h, -d, 1, 0), (-w, -h, d, 0, 1)]
    idx = np.array([0, 1, 2, 3, 1, 0, 4, 5, 6, 6, 5, 7, 8, 9, 10, 11, 9, 8, 12, 13, 14, 14, 13, 15,
                    16, 17, 18, 18, 17, 19, 20, 21, 22, 23, 21, 20], dtype=np.uint32).reshape(-1, 3)
    pos = np.array([x[:3] for x in v], dtype=f32)
    uv = np.array([x[3:] for x in v], dtype=f32)
    nrm = np.zeros_like(pos)
    for a, b, c in idx:                       # calculateNormal (mesh.go:2171-2179), Cross (vector.go:164-180)
        v0 = pos[b] - pos[a]
        v1 = pos[c] - pos[b]
        cr = np.array([v0[1] * v1[2] - v1[1] * v0[2], v0[2] * v1[0] - v1[2] * v0[0], v0[0] * v1[1] - v1[0] * v0[1]],
                      dtype=f32)
        l = f32(np.sqrt(np.float64(cr[0] * cr[0] + cr[1] * cr[1] + cr[2] * cr[2])))
        if not (l < f32(1e-8) or l == f32(1)):
            cr = (cr / l).astype(f32)
        for k in (a, b, c):
            nrm[k] = cr
    mesh = Mesh("Cube", pos, nrm, idx, uvs=uv)
    mesh.mesh_parts[0].material = Material("Cube")
    return mesh


def value_noise(rng, xs, zs, cell, octaves=3):
    """Smooth value noise in [0, 1) on a grid of sample coordinates."""
    total = np.zeros((len(zs), len(xs)), dtype=np.float64)
    amp, norm = 1.0, 0.0
    for o in range(octaves):
        c = cell / (2 ** o)
        gx = xs / c
        gz = zs / c
        nx, nz = int(np.floor(gx.max())) + 2, int(np.floor(gz.max())) + 2
        lattice = rng.floats(nx * nz).astype(np.float64).reshape(nz, nx)
        ix, iz = np.floor(gx).astype(np.int64), np.floor(gz).astype(np.int64)
        fx, fz = gx - ix, gz - iz
        sx, sz = fx * fx * (3 - 2 * fx), fz * fz * (3 - 2 * fz)
        a = lattice[np.ix_(iz, ix)]
        b = lattice[np.ix_(iz, ix + 1)]
        cc = lattice[np.ix_(iz + 1, ix)]
        d = lattice[np.ix_(iz + 1, ix + 1)]
        top = a + (b - a) * sx[None, :]
        bot = cc + (d - cc) * sx[None, :]
        total += amp * (top + (bot - top) * sz[:, None])
        norm += amp
        amp *= 0.5
    return total / norm


def grid_mesh(rng, nx, nz, spacing, amplitude, name="grid", colors=True, noise_cell=None):
    """nx*nz heightfield grid with value-noise heights, finite-difference
    normals and (optionally) random vertex colours; faces point +Y."""
    xs = np.arange(nx, dtype=np.float64) * spacing
    zs = np.arange(nz, dtype=np.float64) * spacing
    cell = noise_cell if noise_cell is not None else max(spacing * 4, (nx * spacing) / 4)
    h = value_noise(rng, xs, zs, cell) * amplitude
    hx = np.gradient(h, spacing, axis=1)
    hz = np.gradient(h, spacing, axis=0)
    n = np.stack([-hx, np.ones_like(h), -hz], axis=-1)
    n /= np.linalg.norm(n, axis=-1, keepdims=True)
    X, Z = np.meshgrid(xs, zs)
    pos = np.stack([X, h, Z], axis=-1).reshape(-1, 3).astype(f32)
    nrm = n.reshape(-1, 3).astype(f32)
    uv = np.stack([X / max(xs[-1], 1e-9), Z / max(zs[-1], 1e-9)], axis=-1).reshape(-1, 2).astype(f32)
    j, i = np.meshgrid(np.arange(nz - 1), np.arange(nx - 1), indexing="ij")
    v00 = (j * nx + i).ravel()
    v10 = (j * nx + i + 1).ravel()
    v01 = ((j + 1) * nx + i).ravel()
    v11 = ((j + 1) * nx + i + 1).ravel()
    # (x0,z0),(x0,z1),(x1,z0) and (x1,z0),(x0,z1),(x1,z1): (b-a)x(c-a) points +Y
    tris = np.empty((len(v00) * 2, 3), dtype=np.uint32)
    tris[0::2] = np.stack([v00, v01, v10], axis=1)
    tris[1::2] = np.stack([v10, v01, v11], axis=1)
    cols = None
    if colors:
        cols = np.ones((nx * nz, 4), dtype=f32)
        cols[:, :3] = rng.floats(nx * nz * 3).reshape(-1, 3)
    return Mesh(name, pos, nrm, tris, colors=cols, uvs=uv)


# ---------------------------------------------------------------------------
# scenes
# ---------------------------------------------------------------------------

class Workload:
    """One Camera.Render call's inputs: camera, scene, light list, model list."""

    def __init__(self, name, camera, scene, lights, models):
        self.name = name
        self.camera = camera
        self.scene = scene
        self.lights = lights
        self.models = models
        seen, meshes = set(), []
        for m in with_batched(models):   # dynamically batched models draw through their owners
            if m.mesh is not None and id(m.mesh) not in seen:
                seen.add(id(m.mesh))
                meshes.append(m.mesh)
        self.meshes = meshes
        self.volumes = [l for l in lights if isinstance(l, LightVolume)]

    @property
    def animations(self):
        """The distinct animations the models play (read when rendered, so a
        model may be given one after the workload is made)."""
        anims = {}
        for m in with_batched(self.models):
            an = getattr(m, "animation", None)
            if an is not None:
                anims[id(an)] = an
        return list(anims.values())

    @property
    def triangle_count(self):
        return int(sum(m.mesh.triangle_count for m in self.models if m.mesh is not None))

    @property
    def vertex_count(self):
        return int(sum(m.mesh.vertex_count for m in self.models if m.mesh is not None))


def _world(ambient):
    w = World()
    w.ambient_light.energy = f32(ambient)
    return w


def _point(name, pos, rng_range, energy=1.0, color=(1, 1, 1)):
    p = PointLight(name, *color, energy)
    p.set_local_position(*pos)
    p.range = f32(rng_range)
    return p


def _directional(name, direction, energy=1.0, color=(1, 1, 1)):
    d = DirectionalLight(name, *color, energy)
    d.set_local_rotation(rotation_from_forward(direction))
    return d


def icosphere_scene(seed=DEFAULT_SEED, level=3):
    """Single model: 642-vertex icosphere at the origin with a seeded rotation,
    white vertex colour; 640x360 camera at z=+4, FOV 75; ambient 0.2, one
    directional light, one point light at (2,2,2) with range 10."""
    rng = SplitMix64(seed)
    mesh = icosphere(level, 1.0, color=(1, 1, 1, 1))
    model = Model("icosphere", mesh)
    q = uniform_quaternions(rng.child(1), 1)[0]
    model.set_local_rotation(quat_to_matrix(*q))
    cam = Camera("camera", 640, 360)
    cam.fov = f32(75)
    cam.set_local_position(0, 0, 4)
    d = _directional("sun", unit_vectors(rng.child(2), 1)[0])
    p = _point("point", (2, 2, 2), 10)
    return Workload("icosphere_642v", cam, Scene("icosphere", _world(0.2)), [d, p], [model])


def cubes_scene(seed=DEFAULT_SEED, count=4096, box=200.0):
    """4,096 shared unit cubes: positions uniform in a 200^3 box, uniform
    rotations, isotropic scale in [0.5, 2]; 1920x1080 camera at the centre of
    the +Z face looking -Z, FOV 60, far 400; ambient 0.2, one directional and
    8 point lights (range 30)."""
    rng = SplitMix64(seed)
    mesh = cube_mesh(1, 1, 1)
    half = box / 2
    pos = rng.child(1).uniform(-half, half, 3 * count).reshape(count, 3)
    quats = uniform_quaternions(rng.child(2), count)
    scales = rng.child(3).uniform(0.5, 2.0, count)
    models = []
    for i in range(count):
        m = Model("cube%d" % i, mesh)
        m.set_local_position(*pos[i])
        m.set_local_rotation(quat_to_matrix(*quats[i]))
        m.set_local_scale(scales[i], scales[i], scales[i])
        models.append(m)
    cam = Camera("camera", 1920, 1080)
    cam.fov = f32(60)
    cam.far = f32(400)
    cam.set_local_position(0, 0, half)
    lights = [_directional("sun", unit_vectors(rng.child(4), 1)[0])]
    lp = rng.child(5).uniform(-half, half, 24).reshape(8, 3)
    lights += [_point("point%d" % i, lp[i], 30) for i in range(8)]
    return Workload("cubes_4096", cam, Scene("cubes", _world(0.2)), lights, models)


def heightfield_scene(seed=DEFAULT_SEED, n=1024, amplitude=20.0):
    """One 1024x1024-vertex heightfield (2,093,058 triangles, one MeshPart):
    value-noise heights with amplitude 20, random vertex colours,
    finite-difference normals; camera inside the terrain's bounding box at
    height 5; 16 point lights (range 15-60), one directional light and one
    LightVolume (cell 32) over the bounding box."""
    rng = SplitMix64(seed)
    mesh = grid_mesh(rng.child(1), n, n, 1.0, amplitude, name="heightfield", noise_cell=64.0)
    model = Model("terrain", mesh)
    cam = Camera("camera", 1920, 1080)
    cam.fov = f32(75)
    cam.far = f32(1000)
    c = (n - 1) / 2.0
    cam.set_local_position(c, 5.0, c)
    yaw = float(rng.child(2).floats(1)[0]) * 2 * np.pi
    cam.set_local_rotation(axis_angle_matrix((0, 1, 0), yaw))
    lr = rng.child(3)
    lights = [_directional("sun", unit_vectors(lr, 1)[0])]
    lp = np.stack([lr.uniform(0, n - 1, 16), lr.uniform(0, 25, 16), lr.uniform(0, n - 1, 16)], axis=1)
    ranges = lr.uniform(15, 60, 16)
    lights += [_point("point%d" % i, lp[i], ranges[i]) for i in range(16)]
    dmin, dmax = mesh.dimensions
    vol = LightVolume("volume", dmin, dmax, (32, 32, 32))
    vol.cells[:, :3] = lr.floats(len(vol.cells) * 3).reshape(-1, 3)
    lights.append(vol)
    return Workload("heightfield_1024", cam, Scene("terrain", _world(0.2)), lights, [model])


def _mixed_mesh(rng, target_v, kind, idx):
    if kind == "sphere":
        level = _sphere_level(target_v)
        return icosphere(level, 4.0, name="sphere%d" % idx, color=(1, 1, 1, 1))
    nx = max(2, int(round(math.sqrt(target_v))))
    nz = max(2, int(round(target_v / nx)))
    return grid_mesh(rng, nx, nz, 8.0 / (nx - 1), 1.5, name="grid%d" % idx)


def _skin(model, rng, bones=32, max_angle_deg=30.0):
    """32-bone armature: 4 influences per vertex with Dirichlet(1) weights
    (normalised in float32), bone matrices = random rotation of at most 30
    degrees composed with the model's placement (identity inverse-bind)."""
    mesh = model.mesh
    V = mesh.vertex_count
    # four distinct bones per vertex: the first four steps of a Fisher-Yates
    # shuffle of 0..bones-1
    perm = np.tile(np.arange(bones, dtype=np.int64), (V, 1))
    rows = np.arange(V)
    for k in range(4):
        j = k + rng.ints(0, bones - k, V)
        pk, pj = perm[rows, k].copy(), perm[rows, j].copy()
        perm[rows, k], perm[rows, j] = pj, pk
    bidx = perm[:, :4].astype(np.uint16)
    e = -np.log(np.maximum(rng.floats(V * 4).astype(np.float64), 1e-12)).reshape(V, 4)
    w = (e / e.sum(axis=1, keepdims=True)).astype(f32)
    mesh.bone_indices = np.ascontiguousarray(bidx)
    mesh.bone_weights = np.ascontiguousarray(w)
    axes = unit_vectors(rng, bones)
    angles = rng.floats(bones).astype(np.float64) * math.radians(max_angle_deg)
    place = model.transform()
    pal = np.stack([mat_mult(axis_angle_matrix(axes[k], angles[k]), place) for k in range(bones)])
    model.skinned = True
    model.bone_influences = pal.astype(f32)


def _sphere_level(target_v):
    return int(np.clip(round(math.log(max(target_v - 2, 10) / 10.0, 4)), 1, 6))


def _mixed_one(args):
    """Model i of a mixed set; its mesh and skin come from its own streams."""
    i, target, pos, quat, seed, skinned, sphere = args
    base = SplitMix64(seed)
    mesh = sphere if sphere is not None else _mixed_mesh(base.child(16).child(i), target,
                                                         "sphere" if i % 4 == 3 else "grid", i)
    m = Model("model%d" % i, mesh)
    m.set_local_position(*pos)
    m.set_local_rotation(quat_to_matrix(*quat))
    if skinned:
        _skin(m, base.child(17).child(i))
    return m


def _mixed_chunk(items):
    return [_mixed_one(a) for a in items]


def mixed_models(rng, count, skinned_every=4, spread=(60.0, 60.0, 30.0, 250.0), vmin=24, vmax=65536, workers=1,
                 only=None):
    """`count` models with vertex counts log-uniform in [vmin, vmax]: grids
    (with random vertex colours) and, every fourth model, shared-size
    icospheres; one model in `skinned_every` is skinned.  Model i draws its
    mesh and skin from streams of its own, so `workers` processes build the
    same models as one, and `only` (a set of indices) builds just those,
    returned as a dict {index: Model}."""
    sx, sy, z0, z1 = spread
    lv = rng.child(11).floats(count).astype(np.float64)
    targets = np.exp(np.log(vmin) + (np.log(vmax) - np.log(vmin)) * lv)
    pos = np.stack([rng.child(12).uniform(-sx, sx, count), rng.child(13).uniform(-sy, sy, count),
                    rng.child(14).uniform(-z1, -z0, count)], axis=1)
    quats = uniform_quaternions(rng.child(15), count)
    sphere_cache = {}
    items = []
    for i in range(count):
        if only is not None and i not in only:
            continue
        skinned = (i % skinned_every) == 1
        sphere = None
        if i % 4 == 3 and not skinned:        # unskinned icospheres share one mesh per level
            level = _sphere_level(targets[i])
            if level not in sphere_cache:
                sphere_cache[level] = _mixed_mesh(None, targets[i], "sphere", i)
            sphere = sphere_cache[level]
        items.append((i, float(targets[i]), pos[i], quats[i], int(rng.seed), skinned, sphere))
    if only is not None:
        return {it[0]: m for it, m in zip(items, _mixed_chunk(items))}
    if workers <= 1 or count < 64:
        return _mixed_chunk(items)
    import multiprocessing as mp
    work = [it for it in items if it[6] is None]
    chunks = [work[k::workers] for k in range(workers)]
    with mp.get_context("fork").Pool(workers) as pool:
        built = {}
        for res in pool.imap_unordered(_mixed_chunk, chunks):
            for m in res:
                built[int(m.name[5:])] = m
    return [built[it[0]] if it[6] is None else _mixed_one(it) for it in items]


def mixed_scene(seed=DEFAULT_SEED, count=512):
    """512 models of uneven size (V log-uniform in [24, 65536]), 128 skinned
    with 32-bone armatures; ambient 0.2, 8 point lights and one directional
    light; 1920x1080 camera at the origin looking -Z, FOV 60, far 400."""
    rng = SplitMix64(seed)
    models = mixed_models(rng, count)
    cam = Camera("camera", 1920, 1080)
    cam.fov = f32(60)
    cam.far = f32(400)
    lr = rng.child(21)
    lights = [_directional("sun", unit_vectors(lr, 1)[0])]
    lp = np.stack([lr.uniform(-60, 60, 8), lr.uniform(-60, 60, 8), lr.uniform(-250, -30, 8)], axis=1)
    lights += [_point("point%d" % i, lp[i], 80) for i in range(8)]
    return Workload("mixed_512", cam, Scene("mixed", _world(0.2)), lights, models)


def armature_poses(rng, bones, place, max_angle_deg=20.0):
    """An armature of `bones` bones as node state, parents first: bone b > 0
    hangs under a uniformly drawn earlier bone (bone 0 under the armature
    node), local offsets in [-0.1, 0.1), rotations of at most `max_angle_deg`
    about random axes, unit scale, identity inverse bind.  Returns the
    BonePose array and the armature node's transform (`place`)."""
    poses = (A.BonePose * bones)()
    pos = rng.uniform(-0.1, 0.1, 3 * bones).reshape(bones, 3)
    axes = unit_vectors(rng, bones)
    angles = rng.floats(bones).astype(np.float64) * math.radians(max_angle_deg)
    parents = [-1] + [int(rng.ints(0, b, 1)[0]) for b in range(1, bones)]
    eye = A.mat4(identity())
    for b in range(bones):
        p = poses[b]
        p.position[:] = [float(x) for x in pos[b]]
        p.scale[:] = [1.0, 1.0, 1.0]
        p.parent = parents[b]
        p.rotation = A.mat4(axis_angle_matrix(axes[b], angles[b]))
        p.inverse_bind = eye
    return poses, np.asarray(place, dtype=f32).copy()


def random_animation(rng, bones, keys=24, length=1.0, name="clip"):
    """An Animation of `bones` channels over `length` seconds: every bone
    has a rotation track of `keys` keyframes (unit quaternions within 25
    degrees of identity), every third bone a position track of 5 keyframes
    (offsets within 0.05), every fifth bone a scale track of 3 keyframes; one
    bone in seven steps (constant interpolation) and one in eleven has no
    tracks at all."""
    from .scene import Animation
    an = Animation(name, bones)
    times = np.linspace(0, length, keys).astype(f32)
    for b in range(bones):
        if b % 11 == 10:
            continue
        ip = A.INTERPOLATION_CONSTANT if b % 7 == 6 else A.INTERPOLATION_LINEAR
        axes = unit_vectors(rng, keys).astype(np.float64)
        ang = rng.floats(keys).astype(np.float64) * math.radians(25.0)
        q = np.concatenate([axes * np.sin(ang / 2)[:, None], np.cos(ang / 2)[:, None]], axis=1)
        if b % 4 == 3:
            q[1::2] *= -1                                   # sign flips: the shortest-arc branch of Slerp
        an.add_track(b, "rotation", times, q.astype(f32), ip)
        if b % 3 == 0:
            an.add_track(b, "position", np.linspace(0, length, 5).astype(f32),
                         rng.uniform(-0.05, 0.05, 15).reshape(5, 3), ip)
        if b % 5 == 0:
            an.add_track(b, "scale", np.linspace(0, length, 3).astype(f32),
                         rng.uniform(0.9, 1.1, 9).reshape(3, 3), ip)
    return an


def crowd_scene(seed=DEFAULT_SEED, count=256, bones=64, level=4, clips=8):
    """`count` characters sharing one 2,562-vertex icosphere mesh skinned to
    `bones` bones (4 influences), each with its own armature given as poses
    and playing one of `clips` shared one-second animations at its own
    playhead, so keyframe sampling and palettes run on the device every
    frame; a 16-wide grid 4 units apart in front of a 1920x1080 camera (FOV
    60, far 400); ambient 0.2, one directional and four point lights."""
    rng = SplitMix64(seed)
    mesh = icosphere(level, 1.0, name="character")
    anims = [random_animation(rng.child(500 + k), bones, name="clip%d" % k) for k in range(clips)]
    playheads = rng.uniform(-0.1, 1.1, count)          # before, inside and after the clips
    models = []
    for i in range(count):
        m = Model("character%d" % i, mesh)
        m.set_local_position(-30 + 4 * (i % 16), -6 + 4 * ((i // 16) % 4), -40 - 6 * (i // 64))
        q = uniform_quaternions(rng.child(100 + i), 1)[0]
        m.set_local_rotation(quat_to_matrix(*q))
        if i == 0:
            _skin(m, rng.child(7), bones=bones)      # the mesh's bone indices and weights
        m.skinned = True
        m.bone_influences = None
        m.bone_poses, m.armature_parent = armature_poses(rng.child(1000 + i), bones, m.transform())
        if clips:
            m.animation = anims[i % clips]
            m.animation_time = float(playheads[i])
        models.append(m)
    cam = Camera("camera", 1920, 1080)
    cam.fov = f32(60)
    cam.far = f32(400)
    lr = rng.child(21)
    lights = [_directional("sun", unit_vectors(lr, 1)[0])]
    lp = np.stack([lr.uniform(-30, 30, 4), lr.uniform(-8, 8, 4), lr.uniform(-60, -30, 4)], axis=1)
    lights += [_point("point%d" % i, lp[i], 60) for i in range(4)]
    return Workload("crowd_%d" % count, cam, Scene("crowd", _world(0.2)), lights, models)


class Job:
    """One serving request: a camera, a subset of the pool and a light list."""

    def __init__(self, arrival_s, camera, model_ids, lights):
        self.arrival_s = arrival_s
        self.camera = camera
        self.model_ids = model_ids
        self.lights = lights


def serving_trace(seed=DEFAULT_SEED, pool_size=10000, jobs=2000, rate_hz=2000.0, max_models=200, max_lights=16,
                  workers=1, models_for=None):
    """A resident pool of `pool_size` models (the mixed distribution) and a
    Poisson job trace at `rate_hz`: each job has a seeded camera (640x360 to
    1920x1080), 1-`max_models` models and 1-`max_lights` lights.

    models_for: a function of the job list returning the pool indices to
    build; the models then come back as a dict {index: Model} holding just
    those (each identical to the full pool's model of that index)."""
    rng = SplitMix64(seed)
    out = _serving_jobs(rng.child(32), pool_size, jobs, rate_hz, max_models, max_lights)
    only = None if models_for is None else set(int(i) for i in models_for(out))
    models = mixed_models(rng.child(31), pool_size, workers=workers, only=only)
    return models, out


def _serving_jobs(jr, pool_size, jobs, rate_hz, max_models, max_lights):
    gaps = -np.log(np.maximum(jr.floats(jobs).astype(np.float64), 1e-12)) / rate_hz
    arrivals = np.cumsum(gaps)
    res = [(640, 360), (1280, 720), (1600, 900), (1920, 1080)]
    out = []
    for j in range(jobs):
        w, h = res[int(jr.ints(0, len(res), 1)[0])]
        cam = Camera("camera%d" % j, w, h)
        cam.fov = f32(jr.uniform(50, 80, 1)[0])
        cam.far = f32(400)
        n = int(jr.ints(1, max_models + 1, 1)[0])
        ids = np.sort(jr.ints(0, pool_size, n))
        nl = int(jr.ints(1, max_lights + 1, 1)[0])
        lights = [_directional("sun%d" % j, unit_vectors(jr, 1)[0])]
        for k in range(nl - 1):
            p = (jr.uniform(-60, 60, 1)[0], jr.uniform(-60, 60, 1)[0], jr.uniform(-250, -30, 1)[0])
            lights.append(_point("p%d_%d" % (j, k), p, 80))
        out.append(Job(float(arrivals[j]), cam, ids, lights))
    return out


SCENES = {
    "icosphere": icosphere_scene,
    "cubes": cubes_scene,
    "heightfield": heightfield_scene,
    "mixed": mixed_scene,
    "crowd": crowd_scene,
    "icosphere_ref": reference_icosphere_scene,
}
