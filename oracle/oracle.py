"""ctypes binding of the CPU oracle (oracle/liboracle.so) -- test infrastructure.

Only tests/, __graft_entry__.smoke() and the cpu_baseline leg of bench.py
import this module, to check the device or to time a CPU baseline.  See
oracle.h for which reference files each oracle function restates.
"""
from __future__ import annotations

import ctypes as C
import math
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "liboracle.so")

ORC_TRIANGLE, ORC_SPHERE, ORC_CYLINDER, ORC_CONE, ORC_CURVE = 0, 1, 2, 4, 5
ORC_RAYCAST, ORC_AO, ORC_DIRECT = 0, 1, 2
ORC_LIGHT_DIRECTIONAL, ORC_LIGHT_POINT, ORC_LIGHT_SPOT = 0, 1, 2
ORC_COLOR_F32, ORC_COLOR_SRGB8, ORC_COLOR_UNORM8 = 0, 1, 2
# bits of the per-pixel "ambiguous" flag of volume marches: a segment-end or
# slab tie (one sample more or less), an early-termination tie at opacity 0.99
AMB_TIE_T, AMB_TIE_STOP = 1, 2
LEAF = 0x80000000

vp = C.c_void_p


class orc_bvh(C.Structure):
    _fields_ = [("n", C.c_uint32), ("order", vp), ("child", vp), ("node_boxes", vp),
                ("root_box", C.c_float * 6), ("leaf_boxes", vp)]


class orc_geom(C.Structure):
    _fields_ = [("type", C.c_int32), ("num_prims", C.c_uint32), ("position", vp), ("index", vp),
                ("radius", vp), ("global_radius", C.c_float), ("caps", C.c_int32),
                ("primitive_id", vp), ("vertex_normal", vp), ("cull_backfaces", C.c_int32),
                ("color", C.c_float * 3), ("opacity", C.c_float), ("surface_id", C.c_uint32)]


class orc_entry(C.Structure):
    _fields_ = [("xfm", C.c_float * 12), ("id", C.c_uint32), ("geom", C.c_uint32)]


class orc_volume(C.Structure):
    _fields_ = [("dims", C.c_uint32 * 3), ("origin", C.c_float * 3), ("spacing", C.c_float * 3),
                ("data", vp), ("lut", vp), ("value_range", C.c_float * 2),
                ("unit_distance", C.c_float), ("id", C.c_uint32), ("inst_id", C.c_uint32)]


_ENTRY = np.dtype([("xfm", np.float32, 12), ("id", np.uint32), ("geom", np.uint32)])


class orc_light(C.Structure):
    _fields_ = [("type", C.c_int32), ("radiance", C.c_float * 3), ("pos", C.c_float * 3),
                ("dir", C.c_float * 3), ("cos_outer", C.c_float), ("cos_inner", C.c_float)]


class orc_scene(C.Structure):
    _fields_ = [("num_geoms", C.c_uint32), ("geoms", vp), ("bvhs", vp),
                ("num_entries", C.c_uint32), ("entries", vp), ("num_volumes", C.c_uint32),
                ("volumes", vp), ("num_lights", C.c_uint32), ("lights", vp), ("tlas", vp)]


class orc_camera(C.Structure):
    _fields_ = [("pos", C.c_float * 3), ("dir", C.c_float * 3), ("up", C.c_float * 3),
                ("fovy", C.c_float), ("aspect", C.c_float), ("region", C.c_float * 4)]


class orc_renderer(C.Structure):
    _fields_ = [("kind", C.c_int32), ("background", C.c_float * 4),
                ("ambient_color", C.c_float * 3), ("ambient_radiance", C.c_float),
                ("pixel_samples", C.c_int32), ("ambient_samples", C.c_int32),
                ("ao_distance", C.c_float), ("cull_backfaces", C.c_int32),
                ("volume_sampling_rate", C.c_float), ("volume_jitter", C.c_int32)]


class orc_frame(C.Structure):
    _fields_ = [("width", C.c_uint32), ("height", C.c_uint32), ("color_format", C.c_int32),
                ("color_f32", vp), ("color_u32", vp), ("depth", vp), ("prim_id", vp),
                ("obj_id", vp), ("inst_id", vp), ("frame_id", C.c_uint32), ("stats", vp),
                ("threads", C.c_int32), ("ambiguous", vp)]


class orc_hit(C.Structure):
    _fields_ = [("t", C.c_float), ("entry", C.c_uint32), ("prim", C.c_uint32),
                ("ng", C.c_float * 3), ("ns", C.c_float * 3), ("u", C.c_float), ("v", C.c_float)]


_lib = None


def lib() -> C.CDLL:
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError("oracle/liboracle.so is not built (make -C oracle)")
        L = C.CDLL(LIB_PATH)
        L.orc_pcg_next.restype = C.c_uint32
        L.orc_pcg_float.restype = C.c_float
        L.orc_expand_bits10.restype = C.c_uint32
        L.orc_expand_bits10.argtypes = [C.c_uint32]
        L.orc_bvh_depth.restype = C.c_uint32
        L.orc_sample_field.restype = C.c_double
        L.orc_volume_march.restype = C.c_double
        L.orc_intersect_triangle_f64.restype = C.c_double
        L.orc_pack_unorm4x8.restype = C.c_uint32
        L.orc_diffvol_loss.restype = C.c_double
        L.orc_morton.argtypes = [vp, C.c_uint32, vp, vp]
        L.orc_sort.argtypes = [vp, C.c_uint32, vp, vp]
        L.orc_karras.argtypes = [vp, C.c_uint32, vp, vp]
        L.orc_refit.argtypes = [vp, vp, vp, C.c_uint32, vp, vp]
        L.orc_scene_build.argtypes = [vp]
        L.orc_scene_free.argtypes = [vp]
        L.orc_render.argtypes = [vp, vp, vp, vp]
        L.orc_bvh_depth.argtypes = [vp]
        L.orc_pack_unorm4x8.argtypes = [vp]
        L.orc_linear_to_srgb.argtypes = [vp, vp]
        L.orc_intersect_triangle.argtypes = [vp, vp, vp, vp, vp, vp, vp, vp]
        L.orc_intersect_cone.argtypes = [vp, vp, vp, vp, C.c_float, C.c_float, vp, vp, vp]
        L.orc_intersect_round_cone.argtypes = [vp, vp, vp, vp, C.c_float, C.c_float, vp, vp]
        L.orc_intersect_triangle_f64.argtypes = [vp, vp, vp, vp, vp]
        L.orc_sample_field.argtypes = [vp, vp]
        L.orc_classify.argtypes = [vp, C.c_double, vp]
        L.orc_volume_march.argtypes = [vp, vp, vp, C.c_double, C.c_double, C.c_double, C.c_int,
                                       vp, vp, vp]
        L.orc_pcg_seed.argtypes = [vp, C.c_uint64, C.c_uint64]
        L.orc_pcg_next.argtypes = [vp]
        L.orc_pcg_float.argtypes = [vp]
        L.orc_bounds.argtypes = [C.c_int, C.c_uint32, vp, vp, vp, C.c_float, vp]
        L.orc_bvh_build.argtypes = [C.c_int, C.c_uint32, vp, vp, vp, C.c_float, vp]
        L.orc_trace_closest.argtypes = [vp, vp, vp, C.c_float, C.c_float, C.c_int, vp, vp]
        L.orc_trace_brute.argtypes = [vp, vp, vp, C.c_float, C.c_float, C.c_int, vp]
        L.orc_trace_any.argtypes = [vp, vp, vp, C.c_float, C.c_float, vp]
        L.orc_majorant_grid.argtypes = [vp, vp]
        L.orc_majorant_grid.restype = None
        L.orc_sample_distance.argtypes = [vp, vp, C.c_uint32, vp, vp, vp, C.c_uint64, vp, vp,
                                          vp]
        L.orc_sample_distance.restype = None
        L.orc_surface_attenuation.argtypes = [vp, vp, vp, C.c_float, C.c_float]
        L.orc_surface_attenuation.restype = C.c_float
        L.orc_trace_rays.argtypes = [vp, C.c_uint32, vp, vp, C.c_float, C.c_float, vp, vp, vp, vp]
        L.orc_intersect_sphere.argtypes = [vp, vp, vp, C.c_float, vp]
        L.orc_intersect_cylinder.argtypes = [vp, vp, vp, vp, C.c_float, C.c_int, vp, vp, vp]
        L.orc_camera_ray.argtypes = [vp, C.c_uint32, C.c_uint32, C.c_float, C.c_float, vp, vp]
        L.orc_volume_image.argtypes = [vp, vp, C.c_uint32, C.c_uint32, C.c_double, C.c_int, vp,
                                       vp, C.c_int32, vp]
        L.orc_tf_discretize.argtypes = [vp, C.c_uint32, C.c_uint32, vp, C.c_uint32, vp,
                                        C.c_float, vp, vp]
        L.orc_diffvol_forward.argtypes = [vp, vp, vp, vp, vp, vp, C.c_double, C.c_double,
                                          C.c_uint32, vp, vp, vp, C.c_int, C.c_int32, vp]
        L.orc_diffvol_loss.argtypes = [vp, vp, vp, vp, vp, vp, C.c_double, C.c_double,
                                       C.c_uint32, vp, vp, vp, C.c_int, C.c_int32]
        _lib = L
    return _lib


def _p(a):
    return None if a is None else a.ctypes.data


# ---------------------------------------------------------------------------
# LBVH pieces
# ---------------------------------------------------------------------------
def bounds(type_: int, pos, index=None, radius=None, global_radius=0.0):
    n = _num_prims(type_, pos, index)
    out = np.empty((n, 6), np.float32)
    lib().orc_bounds(type_, n, _p(pos), _p(index), _p(radius), global_radius, out.ctypes.data)
    return out


def _num_prims(type_, pos, index):
    if type_ == ORC_TRIANGLE:
        return len(index) if index is not None else len(pos) // 3
    if type_ in (ORC_SPHERE, ORC_CURVE):
        return len(index) if index is not None else len(pos)
    return len(index) if index is not None else len(pos) // 2


def morton(boxes: np.ndarray):
    n = len(boxes)
    codes = np.empty(n, np.uint32)
    cbox = np.empty(6, np.float32)
    lib().orc_morton(boxes.ctypes.data, C.c_uint32(n), codes.ctypes.data, cbox.ctypes.data)
    return codes, cbox


def sort(codes: np.ndarray):
    n = len(codes)
    sc = np.empty(n, np.uint32)
    order = np.empty(n, np.uint32)
    lib().orc_sort(codes.ctypes.data, C.c_uint32(n), sc.ctypes.data, order.ctypes.data)
    return sc, order


def karras(sorted_codes: np.ndarray):
    n = len(sorted_codes)
    child = np.zeros(max(2 * (n - 1), 2), np.uint32)
    parent = np.zeros(max(2 * n - 1, 1), np.uint32)
    lib().orc_karras(sorted_codes.ctypes.data, C.c_uint32(n), child.ctypes.data,
                     parent.ctypes.data)
    return child[: 2 * (n - 1)].reshape(-1, 2), parent


def refit(boxes, order, child):
    n = len(boxes)
    nb = np.zeros((max(n - 1, 1), 12), np.float32)
    root = np.zeros(6, np.float32)
    ch = np.ascontiguousarray(child, np.uint32).reshape(-1)
    if ch.size == 0:
        ch = np.zeros(2, np.uint32)
    lib().orc_refit(boxes.ctypes.data, order.ctypes.data, ch.ctypes.data, C.c_uint32(n),
                    nb.ctypes.data, root.ctypes.data)
    return nb[: max(n - 1, 0)], root


def expand_bits10(v: int) -> int:
    return lib().orc_expand_bits10(v)


# ---------------------------------------------------------------------------
# scenes
# ---------------------------------------------------------------------------
_TYPES = {"triangle": ORC_TRIANGLE, "sphere": ORC_SPHERE, "cylinder": ORC_CYLINDER,
          "cone": ORC_CONE, "curve": ORC_CURVE}
IDENTITY12 = np.array([1, 0, 0, 0, 1, 0, 0, 0, 1, 0, 0, 0], np.float32)


def tf_discretize(color, opacity, value_range=(0.0, 1.0), uniform_color=(1, 1, 1, 1),
                  uniform_opacity=1.0):
    lut = np.empty((256, 4), np.float32)
    uc = np.asarray(uniform_color, np.float32)
    vr = np.asarray(value_range, np.float32)
    col = None if color is None else np.ascontiguousarray(color, np.float32)
    op = None if opacity is None else np.ascontiguousarray(opacity, np.float32)
    lib().orc_tf_discretize(_p(col), 0 if col is None else len(col),
                            0 if col is None else col.shape[1], _p(op),
                            0 if op is None else len(op), uc.ctypes.data,
                            C.c_float(uniform_opacity), vr.ctypes.data, lut.ctypes.data)
    return lut


def light_record(L, xfm=None):
    """World-space light record of a scenes.LightDesc under a column-major 3x4
    instance transform, in fp32 with the device host's operation order
    (scene/light/{Directional,Point,Spot}.cpp commit rules; gpu/sampleLight.h
    xfmVec / xfmPoint)."""
    f = np.float32
    m = IDENTITY12 if xfm is None else np.asarray(xfm, np.float32)
    rec = orc_light()
    col = [f(c) for c in L.color]

    def vec(v):
        return [f(f(f(m[i] * v[0]) + f(m[3 + i] * v[1])) + f(m[6 + i] * v[2])) for i in range(3)]

    if L.type == "directional":
        rec.type = ORC_LIGHT_DIRECTIONAL
        d = [f(x) for x in L.direction]
        inv = f(1.0) / np.sqrt(f(f(f(d[0] * d[0]) + f(d[1] * d[1])) + f(d[2] * d[2])))
        nd = [-f(x * inv) for x in d]
        rec.dir[:] = vec(nd)
        irr = f(max(float(L.irradiance), 0.0))
        rec.radiance[:] = [f(c * irr) for c in col]
    else:
        p = [f(x) for x in L.position]
        rec.pos[:] = [f(v + m[9 + i]) for i, v in enumerate(vec(p))]
        inten = f(max(float(L.intensity), 0.0))
        rec.radiance[:] = [f(c * inten) for c in col]
        if L.type == "point":
            rec.type = ORC_LIGHT_POINT
        elif L.type == "spot":
            rec.type = ORC_LIGHT_SPOT
            rec.dir[:] = [f(x) for x in L.direction]
            opening = f(L.opening_angle)
            inner = f(opening - f(f(2.0) * f(L.falloff_angle)))
            rec.cos_outer = float(f(math.cos(float(opening))))
            rec.cos_inner = float(f(math.cos(float(inner))))
        else:
            raise ValueError(L.type)
    return rec


class OracleScene:
    """An orc_scene built from a visrtx_amd.scenes.SceneDesc (same arrays)."""

    def __init__(self, sc):
        self._keep = []
        geoms = (orc_geom * max(len(sc.geoms), 1))()
        for i, g in enumerate(sc.geoms):
            G = geoms[i]
            G.type = _TYPES[g.type]
            G.num_prims = g.num_prims
            pos = np.ascontiguousarray(g.position, np.float32)
            self._keep.append(pos)
            G.position = pos.ctypes.data
            index = g.index
            if g.type == "curve" and index is None:   # Curve.cpp:107-118
                index = np.arange(len(pos) // 2, dtype=np.uint32)
            for name, arr, dt in [("index", index, np.uint32), ("radius", g.radius, np.float32),
                                  ("primitive_id", g.primitive_id, np.uint32),
                                  ("vertex_normal", g.vertex_normal, np.float32)]:
                if arr is not None:
                    a = np.ascontiguousarray(arr, dt)
                    self._keep.append(a)
                    setattr(G, name, a.ctypes.data)
            default_r = 0.01 if g.type == "sphere" else 1.0   # Curve.cpp:49 too
            G.global_radius = g.global_radius if g.global_radius is not None else default_r
            G.caps = 1 if g.caps else 0
            G.cull_backfaces = 1 if g.cull else 0
            G.color[:] = list(g.color)
            G.opacity = g.opacity
            G.surface_id = g.surface_id
        # entries (orc_entry: 12 floats, id, geom) as one structured array:
        # world-level surfaces, then each instance's transforms x its surfaces
        parts = []
        if sc.world_geoms:
            e = np.zeros(len(sc.world_geoms), _ENTRY)
            e["xfm"] = IDENTITY12
            e["id"] = sc.world_id
            e["geom"] = sc.world_geoms
            parts.append(e)
        for inst in sc.instances:
            if not inst.geoms:
                continue
            if inst.xfm is not None and np.ndim(inst.xfm) == 2:
                x = np.asarray(inst.xfm, np.float32).reshape(-1, 12)
                ids = (np.full(len(x), inst.id, np.uint32) if inst.ids is None
                       else np.asarray(inst.ids, np.uint32))
            else:
                x = (IDENTITY12 if inst.xfm is None else np.asarray(inst.xfm, np.float32))[None]
                ids = np.array([inst.id], np.uint32)
            e = np.zeros((len(x), len(inst.geoms)), _ENTRY)
            e["xfm"] = x[:, None, :]
            e["id"] = ids[:, None]
            e["geom"] = np.asarray(inst.geoms, np.uint32)[None, :]
            parts.append(e.reshape(-1))
        ents = np.concatenate(parts) if parts else np.zeros(1, _ENTRY)
        n_ent = sum(len(p) for p in parts)
        entries = ents
        vols = (orc_volume * max(len(sc.volumes), 1))()
        for i, v in enumerate(sc.volumes):
            V = vols[i]
            nz, ny, nx = v.data.shape
            V.dims[:] = [nx, ny, nz]
            V.origin[:] = list(v.origin)
            V.spacing[:] = list(v.spacing)
            d = np.ascontiguousarray(v.data, np.float32)
            lut = tf_discretize(v.color, v.opacity, v.value_range)
            self._keep += [d, lut]
            V.data = d.ctypes.data
            V.lut = lut.ctypes.data
            V.value_range[:] = list(v.value_range)
            V.unit_distance = v.unit_distance
            V.id = v.id
            V.inst_id = sc.world_id
        recs = [light_record(L) for L in getattr(sc, "lights", [])]
        for inst in sc.instances:
            for xfm, _ in inst.transforms():
                recs += [light_record(L, xfm) for L in getattr(inst, "lights", [])]
        lights = (orc_light * max(len(recs), 1))(*recs)
        self._lights = lights
        self._geoms, self._entries, self._vols = geoms, entries, vols
        self.s = orc_scene()
        self.s.num_geoms = len(sc.geoms)
        self.s.geoms = C.addressof(geoms)
        self.s.num_entries = n_ent
        self.s.entries = entries.ctypes.data
        self.s.num_volumes = len(sc.volumes)
        self.s.volumes = C.addressof(vols)
        self.s.num_lights = len(recs)
        self.s.lights = C.addressof(lights)
        if lib().orc_scene_build(C.byref(self.s)) != 0:
            raise RuntimeError("orc_scene_build failed")

    def bvh(self, g: int):
        b = C.cast(self.s.bvhs, C.POINTER(orc_bvh))[g]
        n = b.n
        order = np.ctypeslib.as_array(C.cast(b.order, C.POINTER(C.c_uint32)), (n,)).copy()
        if n > 1:
            child = np.ctypeslib.as_array(C.cast(b.child, C.POINTER(C.c_uint32)),
                                          (2 * (n - 1),)).copy().reshape(-1, 2)
            nb = np.ctypeslib.as_array(C.cast(b.node_boxes, C.POINTER(C.c_float)),
                                       (12 * (n - 1),)).copy().reshape(-1, 12)
        else:
            child = np.zeros((0, 2), np.uint32)
            nb = np.zeros((0, 12), np.float32)
        return order, child, nb, np.array(list(b.root_box), np.float32)

    def depth(self, g: int) -> int:
        b = C.cast(self.s.bvhs, C.POINTER(orc_bvh))
        return lib().orc_bvh_depth(C.byref(b[g]))

    def trace_closest(self, org, dir_, tmin=0.0, tmax=3.4e38, cull=False, stats=None):
        h = orc_hit()
        o = np.asarray(org, np.float32)
        d = np.asarray(dir_, np.float32)
        ok = lib().orc_trace_closest(C.byref(self.s), o.ctypes.data, d.ctypes.data, tmin, tmax,
                                     int(cull), C.byref(h), _p(stats))
        return h if ok else None

    def trace_brute(self, org, dir_, tmin=0.0, tmax=3.4e38, cull=False):
        h = orc_hit()
        o = np.asarray(org, np.float32)
        d = np.asarray(dir_, np.float32)
        ok = lib().orc_trace_brute(C.byref(self.s), o.ctypes.data, d.ctypes.data, tmin, tmax,
                                   int(cull), C.byref(h))
        return h if ok else None

    def trace_any(self, org, dir_, tmin=0.0, tmax=3.4e38) -> bool:
        o = np.asarray(org, np.float32)
        d = np.asarray(dir_, np.float32)
        return bool(lib().orc_trace_any(C.byref(self.s), o.ctypes.data, d.ctypes.data, tmin, tmax,
                                        None))

    def surface_attenuation(self, org, dir_, tmin=0.0, tmax=3.4e38) -> float:
        o = np.asarray(org, np.float32)
        d = np.asarray(dir_, np.float32)
        return float(lib().orc_surface_attenuation(C.byref(self.s), o.ctypes.data, d.ctypes.data,
                                                   tmin, tmax))

    def trace_rays(self, org, dirs, tmin=0.0, tmax=3.4e38):
        n = len(org)
        o = np.ascontiguousarray(org, np.float32)
        d = np.ascontiguousarray(dirs, np.float32)
        t = np.empty(n, np.float32)
        e = np.empty(n, np.uint32)
        p = np.empty(n, np.uint32)
        st = np.zeros(2, np.uint64)
        lib().orc_trace_rays(C.byref(self.s), n, o.ctypes.data, d.ctypes.data, tmin, tmax,
                             t.ctypes.data, e.ctypes.data, p.ctypes.data, st.ctypes.data)
        return t, e, p, st

    def render(self, cam, width, height, kind="raycast", color_format=ORC_COLOR_F32,
               threads=0, **rp):
        r = orc_renderer()
        r.kind = {"ao": ORC_AO, "directLight": ORC_DIRECT, "default": ORC_DIRECT}.get(
            kind, ORC_RAYCAST)
        r.background[:] = list(rp.get("background", (0.0, 0.0, 0.0, 1.0)))
        r.ambient_color[:] = list(rp.get("ambientColor", (1.0, 1.0, 1.0)))
        r.ambient_radiance = rp.get("ambientRadiance", 1.0 if r.kind == ORC_AO else 0.0)
        r.pixel_samples = rp.get("pixelSamples", 1)
        r.ambient_samples = rp.get("ambientSamples", 1)
        r.ao_distance = rp.get("ambientOcclusionDistance", 1e20)
        r.cull_backfaces = int(rp.get("cullTriangleBackfaces", 0))
        r.volume_sampling_rate = min(max(rp.get("volumeSamplingRate", 0.125), 1e-3), 10.0)
        r.volume_jitter = int(bool(rp.get("volumeJitter", False)))
        c = camera(cam)
        f = orc_frame()
        f.width, f.height = width, height
        f.color_format = color_format
        out = {}
        n = width * height
        if color_format == ORC_COLOR_F32:
            out["color"] = np.zeros((height, width, 4), np.float32)
            f.color_f32 = out["color"].ctypes.data
        else:
            out["color"] = np.zeros((height, width), np.uint32)
            f.color_u32 = out["color"].ctypes.data
        out["depth"] = np.zeros((height, width), np.float32)
        out["primitiveId"] = np.zeros((height, width), np.uint32)
        out["objectId"] = np.zeros((height, width), np.uint32)
        out["instanceId"] = np.zeros((height, width), np.uint32)
        stats = np.zeros(4, np.uint64)
        out["ambiguous"] = np.zeros((height, width), np.uint8)
        f.ambiguous = out["ambiguous"].ctypes.data
        f.depth = out["depth"].ctypes.data
        f.prim_id = out["primitiveId"].ctypes.data
        f.obj_id = out["objectId"].ctypes.data
        f.inst_id = out["instanceId"].ctypes.data
        f.stats = stats.ctypes.data
        f.threads = threads
        lib().orc_render(C.byref(self.s), C.byref(c), C.byref(r), C.byref(f))
        out["stats"] = stats
        del n
        return out

    def close(self):
        if self.s is not None:
            lib().orc_scene_free(C.byref(self.s))
            self.s = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def camera(cam) -> orc_camera:
    c = orc_camera()
    c.pos[:] = list(cam.position)
    c.dir[:] = list(cam.direction)
    c.up[:] = list(cam.up)
    c.fovy = cam.fovy
    c.aspect = cam.aspect
    c.region[:] = [0.0, 0.0, 1.0, 1.0]
    return c


def camera_ray(cam, width, height, px, py):
    c = camera(cam)
    o = np.empty(3, np.float32)
    d = np.empty(3, np.float32)
    lib().orc_camera_ray(C.byref(c), width, height, px, py, o.ctypes.data, d.ctypes.data)
    return o, d


def volume_image(vol_desc, cam, width, height, inv_sampling_rate=1.0, early_stop=True, threads=0):
    """FP64 plain volume integral (C.rgb, O) per pixel, plus the sample count."""
    v = orc_volume()
    nz, ny, nx = vol_desc.data.shape
    v.dims[:] = [nx, ny, nz]
    v.origin[:] = list(vol_desc.origin)
    v.spacing[:] = list(vol_desc.spacing)
    d = np.ascontiguousarray(vol_desc.data, np.float32)
    lut = tf_discretize(vol_desc.color, vol_desc.opacity, vol_desc.value_range)
    v.data = d.ctypes.data
    v.lut = lut.ctypes.data
    v.value_range[:] = list(vol_desc.value_range)
    v.unit_distance = vol_desc.unit_distance
    out = np.zeros((height, width, 4), np.float64)
    amb = np.zeros((height, width), np.uint8)
    ns = C.c_uint64()
    c = camera(cam)
    lib().orc_volume_image(C.byref(v), C.byref(c), width, height, inv_sampling_rate,
                           int(early_stop), out.ctypes.data, C.byref(ns), threads,
                           amb.ctypes.data)
    volume_image.last_ambiguous = amb
    return out, ns.value


def _volume_struct(vol_desc):
    """orc_volume of a scenes.VolumeDesc plus the arrays it points at."""
    v = orc_volume()
    nz, ny, nx = vol_desc.data.shape
    v.dims[:] = [nx, ny, nz]
    v.origin[:] = list(vol_desc.origin)
    v.spacing[:] = list(vol_desc.spacing)
    d = np.ascontiguousarray(vol_desc.data, np.float32)
    lut = tf_discretize(vol_desc.color, vol_desc.opacity, vol_desc.value_range)
    v.data = d.ctypes.data
    v.lut = lut.ctypes.data
    v.value_range[:] = list(vol_desc.value_range)
    v.unit_distance = vol_desc.unit_distance
    return v, (d, lut)


def classify(vol_desc, s: float) -> np.ndarray:
    """FP64 RGBA of the volume's transfer function at field value s
    (classifySample, volumeIntegration.h:46-62, position() gpu_math.h:185-189)."""
    v, keep = _volume_struct(vol_desc)
    out = np.zeros(4, np.float64)
    lib().orc_classify(C.byref(v), float(s), out.ctypes.data)
    return out


def majorant_grid(vol_desc) -> np.ndarray:
    """(mz, my, mx) float32 macro-cell majorants (16^3 voxels per cell)."""
    v, keep = _volume_struct(vol_desc)
    m = [(int(x) + 15) // 16 for x in v.dims]
    out = np.zeros((m[2], m[1], m[0]), np.float32)
    lib().orc_majorant_grid(C.byref(v), out.ctypes.data)
    return out


def sample_distance(vol_desc, org, dirs, tminmax, seed, majorant=None):
    """Woodcock tracking per ray: (t, albedo (n, 3), tr)."""
    v, keep = _volume_struct(vol_desc)
    mg = majorant_grid(vol_desc) if majorant is None else np.ascontiguousarray(majorant,
                                                                                np.float32)
    o = np.ascontiguousarray(org, np.float32)
    d = np.ascontiguousarray(dirs, np.float32)
    tm = np.ascontiguousarray(tminmax, np.float32)
    n = len(o)
    t = np.empty(n, np.float32)
    alb = np.empty((n, 3), np.float32)
    tr = np.empty(n, np.float32)
    lib().orc_sample_distance(C.byref(v), mg.ctypes.data, n, o.ctypes.data, d.ctypes.data,
                              tm.ctypes.data, C.c_uint64(seed), t.ctypes.data, alb.ctypes.data,
                              tr.ctypes.data)
    return t, alb, tr


def pack_unorm4x8(v) -> int:
    a = np.asarray(v, np.float32)
    return lib().orc_pack_unorm4x8(a.ctypes.data)


def linear_to_srgb(v):
    a = np.asarray(v, np.float32)
    o = np.empty(4, np.float32)
    lib().orc_linear_to_srgb(a.ctypes.data, o.ctypes.data)
    return o


# ---------------------------------------------------------------------------
# differentiable volume (FP64)
# ---------------------------------------------------------------------------
def _dv_args(dims, origin, spacing, density, lut, value_range):
    d = np.asarray(dims, np.uint32)
    o = np.asarray(origin, np.float32)
    sp = np.asarray(spacing, np.float32)
    den = np.ascontiguousarray(density, np.float64)
    lt = np.ascontiguousarray(lut, np.float64)
    vr = np.asarray(value_range, np.float64)
    return d, o, sp, den, lt, vr


def diffvol_forward(dims, origin, spacing, density, lut, value_range, unit_distance, step, org,
                    dirs, early_stop=False, threads=0):
    """FP64 (C.rgb, O) per ray and the fp32 decision-boundary mask."""
    d, o, sp, den, lt, vr = _dv_args(dims, origin, spacing, density, lut, value_range)
    org = np.ascontiguousarray(org, np.float32)
    dirs = np.ascontiguousarray(dirs, np.float32)
    n = len(org)
    out = np.zeros((n, 4), np.float64)
    amb = np.zeros(n, np.uint8)
    lib().orc_diffvol_forward(d.ctypes.data, o.ctypes.data, sp.ctypes.data, den.ctypes.data,
                              lt.ctypes.data, vr.ctypes.data, unit_distance, step, n,
                              org.ctypes.data, dirs.ctypes.data, out.ctypes.data,
                              int(early_stop), threads, amb.ctypes.data)
    return out, amb


def diffvol_loss(dims, origin, spacing, density, lut, value_range, unit_distance, step, org, dirs,
                 weights, early_stop=False, threads=0):
    """FP64 loss sum(weights * rgba) over the given rays."""
    d, o, sp, den, lt, vr = _dv_args(dims, origin, spacing, density, lut, value_range)
    org = np.ascontiguousarray(org, np.float32)
    dirs = np.ascontiguousarray(dirs, np.float32)
    w = np.ascontiguousarray(weights, np.float64)
    return lib().orc_diffvol_loss(d.ctypes.data, o.ctypes.data, sp.ctypes.data, den.ctypes.data,
                                  lt.ctypes.data, vr.ctypes.data, unit_distance, step, len(org),
                                  org.ctypes.data, dirs.ctypes.data, w.ctypes.data,
                                  int(early_stop), threads)
