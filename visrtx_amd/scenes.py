"""Seeded synthetic scenes and their ANARI construction.

The scene descriptions here are plain numpy data; ``build_anari`` turns one
into ANARI objects on a device, following the object graph of ANARI client
code such as ``examples/simple/testApp_spheres.cpp:167-263`` in the VisRTX
tree (geometry -> surface -> world, camera, renderer, frame).  The
generators are the native PCG32 / value-noise routines of the HIP library
(host code).
"""
from __future__ import annotations

import ctypes as C
import math
from dataclasses import dataclass, field

import numpy as np

from . import _native
from . import anari as A


# ---------------------------------------------------------------------------
# descriptions
# ---------------------------------------------------------------------------
@dataclass
class GeomDesc:
    type: str                       # 'triangle' | 'quad' | 'sphere' | 'cylinder' | 'cone' | 'curve'
    position: np.ndarray            # (V, 3) float32
    index: np.ndarray | None = None
    radius: np.ndarray | None = None
    global_radius: float | None = None
    caps: bool = False
    primitive_id: np.ndarray | None = None
    vertex_normal: np.ndarray | None = None
    color: tuple = (0.8, 0.8, 0.8)
    opacity: float = 1.0
    surface_id: int = 0xFFFFFFFF
    cull: bool = False

    @property
    def num_prims(self) -> int:
        if self.type == "triangle":
            return len(self.index) if self.index is not None else len(self.position) // 3
        if self.type == "quad":   # two triangles per quad
            return 2 * (len(self.index) if self.index is not None else len(self.position) // 4)
        if self.type == "sphere":
            return len(self.index) if self.index is not None else len(self.position)
        if self.type == "curve":    # segment i: vertices index[i], index[i] + 1
            return len(self.index) if self.index is not None else len(self.position) // 2
        return len(self.index) if self.index is not None else len(self.position) // 2


@dataclass
class LightDesc:
    """A directional, point or spot light (scene/light/*.cpp parameters)."""
    type: str                       # "directional" | "point" | "spot"
    color: tuple = (1.0, 1.0, 1.0)
    direction: tuple = (0.0, 0.0, -1.0)
    position: tuple = (0.0, 0.0, 0.0)
    irradiance: float = 1.0         # directional
    intensity: float = 1.0          # point / spot
    opening_angle: float = math.pi  # spot
    falloff_angle: float = 0.1      # spot


@dataclass
class InstanceDesc:
    geoms: list
    xfm: np.ndarray | None = None   # 12 floats, column-major 3x4; or (n, 12): a transform array
    id: int = 0xFFFFFFFF
    lights: list = field(default_factory=list)   # LightDesc in the instance's group
    ids: np.ndarray | None = None   # with a transform array: one id per transform

    def transforms(self):
        """(xfm or None, id) per instanced copy (Instance.cpp:115-128)."""
        if self.xfm is None or np.ndim(self.xfm) == 1:
            return [(self.xfm, self.id)]
        x = np.asarray(self.xfm, np.float32).reshape(-1, 12)
        ids = [self.id] * len(x) if self.ids is None else [int(v) for v in self.ids]
        return list(zip(x, ids))


@dataclass
class VolumeDesc:
    data: np.ndarray                # (nz, ny, nx) float32, x fastest
    origin: tuple = (0.0, 0.0, 0.0)
    spacing: tuple = (1.0, 1.0, 1.0)
    color: np.ndarray | None = None  # (n, 3) or (n, 4)
    opacity: np.ndarray | None = None
    value_range: tuple = (0.0, 1.0)
    unit_distance: float = 1.0
    id: int = 0xFFFFFFFF


@dataclass
class CameraDesc:
    position: tuple
    direction: tuple
    up: tuple
    fovy: float                     # radians
    aspect: float


@dataclass
class SceneDesc:
    geoms: list = field(default_factory=list)
    world_geoms: list = field(default_factory=list)   # indices into geoms
    instances: list = field(default_factory=list)
    volumes: list = field(default_factory=list)
    world_id: int = 0xFFFFFFFF
    lights: list = field(default_factory=list)   # world-level LightDesc


def look_at(pos, target):
    d = np.asarray(target, np.float64) - np.asarray(pos, np.float64)
    return tuple(float(x) for x in d)


# ---------------------------------------------------------------------------
# generators (native, seeded)
# ---------------------------------------------------------------------------
def gen_spheres(n: int, seed: int = 2, rmin: float = 1e-4, rmax: float = 3e-3):
    L = _native.hip()
    centers = np.empty((n, 3), np.float32)
    radii = np.empty(n, np.float32)
    L.vrx_gen_spheres(seed, n, rmin, rmax, centers.ctypes.data, radii.ctypes.data)
    return centers, radii


def gen_heightfield(n: int = 257, seed: int = 1, amplitude: float = 0.2):
    L = _native.hip()
    verts = np.empty((n * n, 3), np.float32)
    idx = np.empty((2 * (n - 1) * (n - 1), 3), np.uint32)
    L.vrx_gen_heightfield(seed, n, n, amplitude, verts.ctypes.data, idx.ctypes.data)
    return verts, idx


def gen_noise_volume(nx: int, ny: int | None = None, nz: int | None = None, seed: int = 3,
                     octaves: int = 4, base_freq: float = 4.0):
    ny = ny or nx
    nz = nz or nx
    L = _native.hip()
    out = np.empty((nz, ny, nx), np.float32)
    L.vrx_gen_noise_volume(seed, nx, ny, nz, octaves, base_freq, out.ctypes.data)
    return out


def gen_sphere_mesh(target_tris: int, seed: int, displacement: float = 0.3):
    L = _native.hip()
    nv, nt = C.c_uint32(), C.c_uint32()
    L.vrx_gen_sphere_mesh_counts(target_tris, C.byref(nv), C.byref(nt))
    verts = np.empty((nv.value, 3), np.float32)
    idx = np.empty((nt.value, 3), np.uint32)
    L.vrx_gen_sphere_mesh(seed, target_tris, displacement, verts.ctypes.data, idx.ctypes.data)
    return verts, idx


# a fixed 5-point colormap (blue, cyan, green, yellow, red)
COLORMAP5 = np.array([[0.0, 0.0, 0.6], [0.0, 0.7, 0.9], [0.1, 0.8, 0.2], [0.95, 0.9, 0.1],
                      [0.9, 0.1, 0.05]], np.float32)


# ---------------------------------------------------------------------------
# the benchmark workloads
# ---------------------------------------------------------------------------
def heightfield_workload(n: int = 257, width: int = 1024, height: int = 768, seed: int = 1):
    """257x257 heightfield, z = 0.2 * 4-octave value noise: 131,072 triangles."""
    verts, idx = gen_heightfield(n, seed)
    sc = SceneDesc(geoms=[GeomDesc("triangle", verts, idx, surface_id=0)], world_geoms=[0],
                   world_id=0)
    cam = CameraDesc((0.5, -1.5, 1.2), look_at((0.5, -1.5, 1.2), (0.5, 0.5, 0.0)), (0, 0, 1),
                     math.radians(45.0), width / height)
    return sc, cam


def spheres_workload(n: int = 10_000_000, width: int = 1920, height: int = 1080, seed: int = 2):
    """n spheres, centers U[0,1]^3, radii log-uniform in [1e-4, 3e-3]."""
    centers, radii = gen_spheres(n, seed)
    sc = SceneDesc(geoms=[GeomDesc("sphere", centers, radius=radii, surface_id=0)],
                   world_geoms=[0], world_id=0)
    cam = CameraDesc((0.5, 0.5, -1.2), (0.0, 0.0, 1.0), (0, 1, 0), math.radians(45.0),
                     width / height)
    return sc, cam


def volume_workload(n: int = 512, width: int = 1920, height: int = 1080, seed: int = 3,
                    data: np.ndarray | None = None):
    """n^3 fp32 fractal noise in [0,1], bounds [0,1]^3, linear opacity ramp 0..0.05."""
    if data is None:
        data = gen_noise_volume(n, seed=seed)
    sp = 1.0 / (n - 1)
    vol = VolumeDesc(data=data, origin=(0.0, 0.0, 0.0), spacing=(sp, sp, sp),
                     color=COLORMAP5, opacity=np.array([0.0, 0.05], np.float32),
                     value_range=(0.0, 1.0), unit_distance=sp, id=0)
    sc = SceneDesc(volumes=[vol], world_id=0)
    pos = (1.9, 1.5, -1.0)
    cam = CameraDesc(pos, look_at(pos, (0.5, 0.5, 0.5)), (0, 1, 0), math.radians(35.0),
                     width / height)
    return sc, cam


def sparse_volume_workload(n: int = 512, blobs: int = 12, width: int = 1920,
                           height: int = 1080, seed: int = 5):
    """n^3 field, bounds [0,1]^3: value 0 except inside ``blobs`` balls
    (radius 0.06..0.14), where it is a smooth bump times fractal noise.  The
    transfer function's opacity is exactly 0 below value 1/3, so most 16^3
    macro cells are empty."""
    rng = np.random.default_rng(seed)
    noise = gen_noise_volume(n, seed=seed)
    ax = np.linspace(0.0, 1.0, n, dtype=np.float32)
    data = np.zeros((n, n, n), np.float32)
    for _ in range(blobs):
        c = rng.uniform(0.15, 0.85, 3)
        r = rng.uniform(0.06, 0.14)
        i0 = [max(int((c[a] - r) * (n - 1)) - 1, 0) for a in range(3)]
        i1 = [min(int((c[a] + r) * (n - 1)) + 2, n) for a in range(3)]
        z, y, x = np.meshgrid(ax[i0[2]:i1[2]], ax[i0[1]:i1[1]], ax[i0[0]:i1[0]], indexing="ij")
        q = ((x - c[0]) ** 2 + (y - c[1]) ** 2 + (z - c[2]) ** 2) / (r * r)
        bump = np.where(q < 1.0, (1.0 - q) ** 2, 0.0).astype(np.float32)
        sl = (slice(i0[2], i1[2]), slice(i0[1], i1[1]), slice(i0[0], i1[0]))
        data[sl] = np.maximum(data[sl], bump * (0.5 + noise[sl]))
    sp = 1.0 / (n - 1)
    vol = VolumeDesc(data=data, origin=(0.0, 0.0, 0.0), spacing=(sp, sp, sp),
                     color=COLORMAP5, opacity=np.array([0.0, 0.0, 0.4, 0.8], np.float32),
                     value_range=(0.0, 1.0), unit_distance=sp * 4, id=0)
    sc = SceneDesc(volumes=[vol], world_id=0)
    pos = (1.9, 1.5, -1.0)
    cam = CameraDesc(pos, look_at(pos, (0.5, 0.5, 0.5)), (0, 1, 0), math.radians(35.0),
                     width / height)
    return sc, cam


def random_rotation(rng) -> np.ndarray:
    """Uniformly distributed rotation (normalised Gaussian quaternion), 3x3 float64."""
    q = rng.normal(size=4)
    w, x, y, z = q / np.linalg.norm(q)
    return np.array([[1 - 2 * (y * y + z * z), 2 * (x * y - z * w), 2 * (x * z + y * w)],
                     [2 * (x * y + z * w), 1 - 2 * (x * x + z * z), 2 * (y * z - x * w)],
                     [2 * (x * z - y * w), 2 * (y * z + x * w), 1 - 2 * (x * x + y * y)]])


def orbit_poses(n: int, center, radius: float, seed: int, fovy_deg: float = 45.0,
                aspect: float = 1.0) -> list:
    """n seeded camera poses on an orbit around ``center``, looking at it."""
    rng = np.random.default_rng(seed)
    c = np.asarray(center, np.float64)
    cams = []
    for i in range(n):
        phi = 2.0 * math.pi * (i + rng.uniform(0.0, 0.5)) / n
        z = rng.uniform(-0.5, 0.5)
        r = math.sqrt(1.0 - z * z)
        pos = tuple(float(v) for v in c + radius * np.array([r * math.cos(phi), z,
                                                              r * math.sin(phi)]))
        cams.append(CameraDesc(pos, look_at(pos, c), (0, 1, 0), math.radians(fovy_deg), aspect))
    return cams


def instanced_spheres_workload(n: int = 1_000_000, seed: int = 21, width: int = 1920,
                               height: int = 1080):
    """One Instance whose 'transform' array holds n elements (Instance.cpp:
    115-128): a unit sphere scaled to radius 2e-3..6e-3 and moved into the
    unit cube, with ids 3k + 7 from an 'id' array."""
    rng = np.random.default_rng(seed)
    s = rng.uniform(2e-3, 6e-3, n).astype(np.float32)
    xf = np.zeros((n, 12), np.float32)
    xf[:, 0] = xf[:, 4] = xf[:, 8] = s
    xf[:, 9:12] = rng.random((n, 3), dtype=np.float32)
    g = GeomDesc("sphere", np.zeros((1, 3), np.float32), radius=np.ones(1, np.float32),
                 surface_id=5)
    inst = InstanceDesc(geoms=[0], xfm=xf, ids=np.arange(n, dtype=np.uint32) * 3 + 7)
    sc = SceneDesc(geoms=[g], instances=[inst], world_id=1)
    cam = CameraDesc((0.5, 0.5, -1.2), (0, 0, 1), (0, 1, 0), math.radians(45), width / height)
    return sc, cam


def batched_workload(views: int = 64, res: int = 512, instances: int = 256, meshes: int = 16,
                     tri_min: int = 10_000, tri_max: int = 2_000_000, box: float = 10.0,
                     seed: int = 9):
    """``meshes`` noise-displaced spheres (radius ~1) with triangle counts
    log-uniform in [tri_min, tri_max], ``instances`` instances of them with
    uniformly random rotations and translations in [0, box]^3, and ``views``
    orbit poses at res x res looking at the box centre."""
    rng = np.random.default_rng(seed)
    counts = np.exp(rng.uniform(math.log(tri_min), math.log(tri_max), meshes)).astype(np.int64)
    geoms = []
    for i, t in enumerate(counts):
        v, idx = gen_sphere_mesh(int(t), seed=seed * 1000 + i)
        geoms.append(GeomDesc("triangle", v, idx, surface_id=i,
                              color=tuple(float(c) for c in rng.uniform(0.2, 0.9, 3))))
    insts = []
    for k in range(instances):
        R = random_rotation(rng)
        t = rng.uniform(0.0, box, 3)
        M = np.zeros(12, np.float32)
        M[0:9] = R.T.reshape(-1)        # column-major 3x3
        M[9:12] = t
        insts.append(InstanceDesc(geoms=[int(rng.integers(meshes))], xfm=M, id=k))
    sc = SceneDesc(geoms=geoms, instances=insts, world_id=0)
    c = 0.5 * box
    cams = orbit_poses(views, (c, c, c), 1.9 * box, seed + 1)
    return sc, cams


# ---------------------------------------------------------------------------
# ANARI construction
# ---------------------------------------------------------------------------
_GEOM_SUB = {"triangle": "triangle", "quad": "quad", "sphere": "sphere", "cylinder": "cylinder",
             "cone": "cone", "curve": "curve"}


def quads_to_triangles(index: np.ndarray | None, num_vertices: int) -> np.ndarray:
    """The two triangles of each quad, as the device splits them
    (scene/surface/geometry/Quad.cpp:124-145): (x, y, w) and (z, w, y) of
    quad (x, y, z, w); a soup uses quad q = vertices 4q .. 4q+3."""
    q = (np.asarray(index, np.uint32).reshape(-1, 4) if index is not None
         else np.arange(num_vertices, dtype=np.uint32).reshape(-1, 4))
    t = np.empty((len(q), 2, 3), np.uint32)
    t[:, 0] = q[:, [0, 1, 3]]
    t[:, 1] = q[:, [2, 3, 1]]
    return t.reshape(-1, 3)


class AnariScene:
    """Handles of one scene on a device (released by ``release``)."""

    def __init__(self, dev: A.Device):
        self.dev = dev
        self.handles = []
        self.geoms = []
        self.world = None

    def keep(self, h):
        self.handles.append(h)
        return h

    def drop(self, h):
        """Release one kept handle now (it is then not released again)."""
        self.handles.remove(h)
        self.dev.release(h)

    def release(self):
        for h in reversed(self.handles):
            self.dev.release(h)
        self.handles.clear()


def _surface(dev: A.Device, s: AnariScene, g: GeomDesc):
    geom = s.keep(dev.new("Geometry", _GEOM_SUB[g.type]))
    s.geoms.append(geom)
    dev.set(geom, "vertex.position", s.keep(dev.array(g.position, A.FLOAT32_VEC3)), A.ARRAY1D)
    if g.index is not None:
        et = {"triangle": A.UINT32_VEC3, "quad": A.UINT32_VEC4, "sphere": A.UINT32,
              "cylinder": A.UINT32_VEC2, "cone": A.UINT32_VEC2, "curve": A.UINT32}[g.type]
        dev.set(geom, "primitive.index", s.keep(dev.array(g.index, et)), A.ARRAY1D)
    if g.radius is not None:
        name = "vertex.radius" if g.type in ("sphere", "cone", "curve") else "primitive.radius"
        dev.set(geom, name, s.keep(dev.array(g.radius, A.FLOAT32)), A.ARRAY1D)
    if g.global_radius is not None:
        dev.set(geom, "radius", g.global_radius, A.FLOAT32)
    if g.type == "cylinder" and g.caps:
        dev.set(geom, "caps", "both", A.STRING)
    if g.primitive_id is not None:
        dev.set(geom, "primitive.id", s.keep(dev.array(g.primitive_id, A.UINT32)), A.ARRAY1D)
    if g.vertex_normal is not None:
        dev.set(geom, "vertex.normal", s.keep(dev.array(g.vertex_normal, A.FLOAT32_VEC3)),
                A.ARRAY1D)
    if g.cull:
        dev.set(geom, "cullBackfaces", 1, A.BOOL)
    dev.commit(geom)
    mat = s.keep(dev.new("Material", "matte"))
    dev.set(mat, "color", g.color, A.FLOAT32_VEC3)
    if g.opacity != 1.0:
        dev.set(mat, "opacity", g.opacity, A.FLOAT32)
        dev.set(mat, "alphaMode", "blend", A.STRING)
    dev.commit(mat)
    surf = s.keep(dev.new("Surface"))
    dev.set(surf, "geometry", geom, A.GEOMETRY)
    dev.set(surf, "material", mat, A.MATERIAL)
    if g.surface_id != 0xFFFFFFFF:
        dev.set(surf, "id", g.surface_id, A.UINT32)
    dev.commit(surf)
    return surf


def _volume(dev: A.Device, s: AnariScene, v: VolumeDesc):
    nz, ny, nx = v.data.shape
    field_ = s.keep(dev.new("SpatialField", "structuredRegular"))
    dev.set(field_, "data", s.keep(dev.array(v.data, A.FLOAT32, dims=(nx, ny, nz))), A.ARRAY3D)
    dev.set(field_, "origin", v.origin, A.FLOAT32_VEC3)
    dev.set(field_, "spacing", v.spacing, A.FLOAT32_VEC3)
    dev.commit(field_)
    vol = s.keep(dev.new("Volume", "transferFunction1D"))
    dev.set(vol, "value", field_, A.SPATIAL_FIELD)
    if v.color is not None:
        et = A.FLOAT32_VEC3 if v.color.shape[1] == 3 else A.FLOAT32_VEC4
        dev.set(vol, "color", s.keep(dev.array(v.color, et)), A.ARRAY1D)
    if v.opacity is not None:
        dev.set(vol, "opacity", s.keep(dev.array(v.opacity, A.FLOAT32)), A.ARRAY1D)
    dev.set(vol, "valueRange", v.value_range, A.FLOAT32_VEC2)
    dev.set(vol, "unitDistance", v.unit_distance, A.FLOAT32)
    if v.id != 0xFFFFFFFF:
        dev.set(vol, "id", v.id, A.UINT32)
    dev.commit(vol)
    return vol


def _light(dev: A.Device, s: AnariScene, L: LightDesc):
    h = s.keep(dev.new("Light", L.type))
    dev.set(h, "color", L.color, A.FLOAT32_VEC3)
    if L.type == "directional":
        dev.set(h, "direction", L.direction, A.FLOAT32_VEC3)
        dev.set(h, "irradiance", L.irradiance, A.FLOAT32)
    else:
        dev.set(h, "position", L.position, A.FLOAT32_VEC3)
        dev.set(h, "intensity", L.intensity, A.FLOAT32)
        if L.type == "spot":
            dev.set(h, "direction", L.direction, A.FLOAT32_VEC3)
            dev.set(h, "openingAngle", L.opening_angle, A.FLOAT32)
            dev.set(h, "falloffAngle", L.falloff_angle, A.FLOAT32)
    dev.commit(h)
    return h


def build_anari(dev: A.Device, sc: SceneDesc) -> AnariScene:
    s = AnariScene(dev)
    surfs = [None] * len(sc.geoms)

    def surf_of(i):
        if surfs[i] is None:
            surfs[i] = _surface(dev, s, sc.geoms[i])
        return surfs[i]

    world = s.keep(dev.new("World"))
    if sc.world_geoms:
        arr = s.keep(dev.array([surf_of(i) for i in sc.world_geoms], A.SURFACE))
        dev.set(world, "surface", arr, A.ARRAY1D)
    if sc.volumes:
        vols = [_volume(dev, s, v) for v in sc.volumes]
        dev.set(world, "volume", s.keep(dev.array(vols, A.VOLUME)), A.ARRAY1D)
    if sc.instances:
        insts = []
        for inst in sc.instances:
            grp = s.keep(dev.new("Group"))
            if inst.geoms:
                arr = s.keep(dev.array([surf_of(i) for i in inst.geoms], A.SURFACE))
                dev.set(grp, "surface", arr, A.ARRAY1D)
            if inst.lights:
                la = s.keep(dev.array([_light(dev, s, L) for L in inst.lights], A.LIGHT))
                dev.set(grp, "light", la, A.ARRAY1D)
            dev.commit(grp)
            ih = s.keep(dev.new("Instance", "transform"))
            dev.set(ih, "group", grp, A.GROUP)
            if inst.xfm is not None and np.ndim(inst.xfm) == 2:
                # a transform array: column-major 4x4 per element
                x = np.asarray(inst.xfm, np.float32).reshape(-1, 4, 3)
                m4 = np.zeros((len(x), 4, 4), np.float32)
                m4[:, :, :3] = x
                m4[:, 3, 3] = 1.0
                dev.set(ih, "transform", s.keep(dev.array(m4, A.FLOAT32_MAT4)), A.ARRAY1D)
                if inst.ids is not None:
                    dev.set(ih, "id", s.keep(dev.array(np.asarray(inst.ids, np.uint32), A.UINT32)),
                            A.ARRAY1D)
            elif inst.xfm is not None:
                dev.set(ih, "transform", np.asarray(inst.xfm, np.float32), A.FLOAT32_MAT3x4)
            if inst.id != 0xFFFFFFFF:
                dev.set(ih, "id", inst.id, A.UINT32)
            dev.commit(ih)
            insts.append(ih)
        dev.set(world, "instance", s.keep(dev.array(insts, A.INSTANCE)), A.ARRAY1D)
    if sc.lights:
        la = s.keep(dev.array([_light(dev, s, L) for L in sc.lights], A.LIGHT))
        dev.set(world, "light", la, A.ARRAY1D)
    if sc.world_id != 0xFFFFFFFF:
        dev.set(world, "id", sc.world_id, A.UINT32)
    dev.commit(world)
    s.world = world
    return s


def make_camera(dev: A.Device, s: AnariScene, cam: CameraDesc):
    c = s.keep(dev.new("Camera", "perspective"))
    dev.set(c, "position", cam.position, A.FLOAT32_VEC3)
    dev.set(c, "direction", cam.direction, A.FLOAT32_VEC3)
    dev.set(c, "up", cam.up, A.FLOAT32_VEC3)
    dev.set(c, "fovy", cam.fovy, A.FLOAT32)
    dev.set(c, "aspect", cam.aspect, A.FLOAT32)
    dev.commit(c)
    return c


def make_renderer(dev: A.Device, s: AnariScene, kind: str = "raycast", **params):
    r = s.keep(dev.new("Renderer", kind))
    types = {"background": A.FLOAT32_VEC4, "ambientColor": A.FLOAT32_VEC3,
             "ambientRadiance": A.FLOAT32, "pixelSamples": A.INT32, "ambientSamples": A.INT32,
             "ambientOcclusionDistance": A.FLOAT32, "cullTriangleBackfaces": A.BOOL,
             "volumeSamplingRate": A.FLOAT32, "sampleLimit": A.INT32, "stackDepthLDS": A.INT32,
             "volumeUnroll": A.INT32, "nodeWidth": A.INT32, "lightFalloff": A.FLOAT32,
             "emptySpaceSkipping": A.BOOL, "volumeJitter": A.BOOL, "longRaySteps": A.INT32}
    for k, v in params.items():
        dev.set(r, k, v, types[k])
    dev.commit(r)
    return r


def make_frame(dev: A.Device, s: AnariScene, world, camera, renderer, width: int, height: int,
               color=A.FLOAT32_VEC4, depth=True, ids=True):
    f = s.keep(dev.new("Frame"))
    dev.set(f, "size", (width, height), A.UINT32_VEC2)
    dev.set(f, "channel.color", color, A.DATA_TYPE)
    if depth:
        dev.set(f, "channel.depth", A.FLOAT32, A.DATA_TYPE)
    if ids:
        dev.set(f, "channel.primitiveId", A.UINT32, A.DATA_TYPE)
        dev.set(f, "channel.objectId", A.UINT32, A.DATA_TYPE)
        dev.set(f, "channel.instanceId", A.UINT32, A.DATA_TYPE)
    dev.set(f, "world", world, A.WORLD)
    dev.set(f, "camera", camera, A.CAMERA)
    dev.set(f, "renderer", renderer, A.RENDERER)
    dev.commit(f)
    return f
