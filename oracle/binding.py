"""ctypes binding of the CPU oracle (liblc_oracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, the smoke check and the benchmark's CPU-baseline
leg, as the checker. The product package never imports this module.
"""
import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "liblc_oracle.so")

HIT_DTYPE = np.dtype([("inst", "<u4"), ("prim", "<u4"), ("u", "<f4"), ("v", "<f4"), ("t", "<f4"), ("pad", "<u4")])
HIT2_DTYPE = np.dtype([("inst", "<u4"), ("prim", "<u4"), ("u", "<f4"), ("v", "<f4"), ("t", "<f4"), ("t2", "<f4")])
COMMITTED_DTYPE = np.dtype([("inst", "<u4"), ("prim", "<u4"), ("u", "<f4"), ("v", "<f4"), ("hit_type", "<u4"),
                            ("t", "<f4")])
BF_DTYPE = np.dtype([("prim", "<u4"), ("pad", "<u4"), ("t", "<f8"), ("u", "<f8"), ("v", "<f8"), ("t2", "<f8")])


class Pcg32(ctypes.Structure):
    _fields_ = [("state", ctypes.c_uint64), ("inc", ctypes.c_uint64)]


class QueryPolicy(ctypes.Structure):
    """orc_query_policy: surface_mode 0 commit / 1 ignore / 2 cutout; procedural_mode 0 ignore / 1 sphere."""
    _fields_ = [("surface_mode", ctypes.c_uint32), ("cutout", ctypes.c_float), ("procedural_mode", ctypes.c_uint32),
                ("terminate_on_first", ctypes.c_uint32)]


class PtParams(ctypes.Structure):
    """orc_pt_params."""
    _fields_ = [(n, ctypes.c_float * 3) for n in ("light_position", "light_u", "light_v", "light_emission",
                                                   "camera_origin", "camera_right", "camera_up", "camera_forward")] + \
               [(n, ctypes.c_uint32) for n in ("width", "height", "spp", "sample_offset", "max_depth", "rr_depth")] + \
               [("seed", ctypes.c_uint64), ("sequence", ctypes.c_uint64)]


PT_MATERIAL_DTYPE = np.dtype([("albedo", "<f4", (3,)), ("emissive", "<u4")])


def build():
    """Compile the oracle in place (make)."""
    subprocess.run(["make", "-s", "-C", _HERE], check=True)


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        L = ctypes.CDLL(_LIB_PATH)
        P, U64, U32, I = ctypes.c_void_p, ctypes.c_uint64, ctypes.c_uint32, ctypes.c_int
        F = ctypes.c_float
        sig = {
            "orc_pcg32_seed": (None, [ctypes.POINTER(Pcg32), U64, U64]),
            "orc_pcg32_next": (U32, [ctypes.POINTER(Pcg32)]),
            "orc_pcg32_advance": (None, [ctypes.POINTER(Pcg32), U64]),
            "orc_pcg32_float": (F, [ctypes.POINTER(Pcg32)]),
            "orc_icosphere_counts": (None, [I, ctypes.POINTER(U32), ctypes.POINTER(U32)]),
            "orc_icosphere": (None, [I, P, P]),
            "orc_rays_orbit_to_ball": (None, [U64, U64, U64, F, P]),
            "orc_soup": (None, [U64, U64, U64, F, F, P, P]),
            "orc_rays_box_sphere": (None, [U64, U64, U64, U64, F, F, P]),
            "orc_instances": (None, [U64, U64, U32, U32, F, F, F, P, P]),
            "orc_batch_sizes": (None, [U64, U64, U64, U32, U32, U32, P]),
            "orc_mesh_build": (P, [P, U64, U64, P, U64, I]),
            "orc_mesh_free": (None, [P]),
            "orc_mesh_node_count": (U64, [P]),
            "orc_scene_create": (P, []),
            "orc_scene_add": (None, [P, P, P, U32]),
            "orc_scene_add_ex": (None, [P, P, P, P, U32, U32]),
            "orc_procedural_create": (P, [P, U64]),
            "orc_procedural_free": (None, [P]),
            "orc_query": (None, [P, P, U64, U32, ctypes.POINTER(QueryPolicy), P, I]),
            "orc_query_candidate": (I, [P, P, U32, ctypes.POINTER(QueryPolicy), U32, U32, U32, ctypes.POINTER(F),
                                        ctypes.POINTER(F), ctypes.POINTER(F)]),
            "orc_scene_commit": (None, [P]),
            "orc_scene_free": (None, [P]),
            "orc_trace_closest": (None, [P, P, U64, U32, P, I]),
            "orc_trace_closest2": (None, [P, P, U64, U32, P, I]),
            "orc_trace_any": (None, [P, P, U64, U32, P, I]),
            "orc_brute_closest_f64": (None, [P, U64, P, U64, P, U64, P, I]),
            "orc_watertight": (I, [P, P, P, P, P, F, F, P, P, P]),
            "orc_pathtrace": (None, [P, P, ctypes.POINTER(PtParams), U64, U64, P, P, I]),
            "orc_sincos_2pi": (None, [F, ctypes.POINTER(F), ctypes.POINTER(F)]),
            "orc_offset_ray_origin": (None, [P, P, P]),
            "orc_pathtrace_rays": (None, [P, P, ctypes.POINTER(PtParams), U64, U64, P, U64, ctypes.POINTER(U64), P, U64,
                                          ctypes.POINTER(U64)]),
        }
        for name, (res, args) in sig.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _lib = L
    return _lib


def _p(a):
    return a.ctypes.data_as(ctypes.c_void_p)


class Rng:
    def __init__(self, initstate, initseq):
        self.s = Pcg32()
        lib().orc_pcg32_seed(ctypes.byref(self.s), initstate, initseq)

    def next(self):
        return lib().orc_pcg32_next(ctypes.byref(self.s))

    def advance(self, delta):
        lib().orc_pcg32_advance(ctypes.byref(self.s), delta)

    def float(self):
        return lib().orc_pcg32_float(ctypes.byref(self.s))


def icosphere(level):
    nv, nt = ctypes.c_uint32(), ctypes.c_uint32()
    lib().orc_icosphere_counts(level, ctypes.byref(nv), ctypes.byref(nt))
    v = np.zeros((nv.value, 4), np.float32)
    t = np.zeros((nt.value, 3), np.uint32)
    lib().orc_icosphere(level, _p(v), _p(t))
    return v, t


def rays_orbit_to_ball(n, initstate=42, initseq=1, orbit=3.0):
    r = np.zeros((n, 8), np.float32)
    lib().orc_rays_orbit_to_ball(initstate, initseq, n, orbit, _p(r))
    return r


def soup(ntris, initstate=42, initseq=2, extent=100.0, sigma=0.5):
    v = np.zeros((3 * ntris, 4), np.float32)
    t = np.zeros((ntris, 3), np.uint32)
    lib().orc_soup(initstate, initseq, ntris, extent, sigma, _p(v), _p(t))
    return v, t


def rays_box_sphere(n, initstate=42, initseq=2, skip=0, lo=0.0, hi=100.0):
    r = np.zeros((n, 8), np.float32)
    lib().orc_rays_box_sphere(initstate, initseq, skip, n, lo, hi, _p(r))
    return r


def instances(count, mesh_count, initstate=42, initseq=4, extent=50.0, slo=0.5, shi=2.0):
    a = np.zeros((count, 12), np.float32)
    m = np.zeros(count, np.uint32)
    lib().orc_instances(initstate, initseq, count, mesh_count, extent, slo, shi, _p(a), _p(m))
    return a, m


def batch_sizes(count, lo, hi, initstate=42, initseq=4, skip=0):
    s = np.zeros(count, np.uint32)
    lib().orc_batch_sizes(initstate, initseq, skip, count, lo, hi, _p(s))
    return s


class Mesh:
    def __init__(self, verts4, tris, nthreads=0):
        self.verts = np.ascontiguousarray(verts4, np.float32)
        self.tris = np.ascontiguousarray(tris, np.uint32)
        stride = self.verts.strides[0]
        self.h = lib().orc_mesh_build(_p(self.verts), stride, self.verts.shape[0], _p(self.tris),
                                      self.tris.shape[0], nthreads)

    def node_count(self):
        return lib().orc_mesh_node_count(self.h)

    def __del__(self):
        if getattr(self, "h", None):
            lib().orc_mesh_free(self.h)
            self.h = None


class Procedural:
    """AABB list (n, 6): min.xyz, max.xyz per primitive."""

    def __init__(self, aabbs):
        self.aabbs = np.ascontiguousarray(aabbs, np.float32).reshape(-1, 6)
        self.h = lib().orc_procedural_create(_p(self.aabbs), self.aabbs.shape[0])

    def __del__(self):
        if getattr(self, "h", None):
            lib().orc_procedural_free(self.h)
            self.h = None


IDENTITY = np.array([1, 0, 0, 0, 0, 1, 0, 0, 0, 0, 1, 0], np.float32)


class Scene:
    """Two-level scene: instances of oracle meshes with row-major 3x4 affines."""

    def __init__(self):
        self.h = lib().orc_scene_create()
        self._keep = []

    def add(self, mesh, affine=IDENTITY, mask=0xFF, opaque=True):
        """mesh: a Mesh or a Procedural."""
        a = np.ascontiguousarray(np.asarray(affine, np.float32).reshape(12))
        self._keep.append((mesh, a))
        if isinstance(mesh, Procedural):
            lib().orc_scene_add_ex(self.h, None, mesh.h, _p(a), mask, 1 if opaque else 0)
        else:
            lib().orc_scene_add_ex(self.h, mesh.h, None, _p(a), mask, 1 if opaque else 0)
        return self

    def query(self, rays, mask=0xFF, surface=0, cutout=0.8, procedural=1, nthreads=0):
        """Nearest committed candidate per ray (COMMITTED_DTYPE)."""
        rays = np.ascontiguousarray(rays, np.float32).reshape(-1, 8)
        out = np.zeros(rays.shape[0], COMMITTED_DTYPE)
        pol = QueryPolicy(surface, cutout, procedural, 0)
        lib().orc_query(self.h, _p(rays), rays.shape[0], mask, ctypes.byref(pol), _p(out), nthreads)
        return out

    def query_candidate(self, ray, inst, prim, hit_type, mask=0xFF, surface=0, cutout=0.8, procedural=1):
        """(t, u, v) if candidate (inst, prim) of kind hit_type is committed for `ray`, else None."""
        ray = np.ascontiguousarray(ray, np.float32).reshape(8)
        pol = QueryPolicy(surface, cutout, procedural, 0)
        t, u, v = ctypes.c_float(), ctypes.c_float(), ctypes.c_float()
        ok = lib().orc_query_candidate(self.h, _p(ray), mask, ctypes.byref(pol), int(inst), int(prim), int(hit_type),
                                       ctypes.byref(t), ctypes.byref(u), ctypes.byref(v))
        return (t.value, u.value, v.value) if ok else None

    def commit(self):
        lib().orc_scene_commit(self.h)
        return self

    def trace_closest(self, rays, mask=0xFF, nthreads=0):
        rays = np.ascontiguousarray(rays, np.float32)
        out = np.zeros(rays.shape[0], HIT_DTYPE)
        lib().orc_trace_closest(self.h, _p(rays), rays.shape[0], mask, _p(out), nthreads)
        return out

    def trace_closest2(self, rays, mask=0xFF, nthreads=0):
        rays = np.ascontiguousarray(rays, np.float32)
        out = np.zeros(rays.shape[0], HIT2_DTYPE)
        lib().orc_trace_closest2(self.h, _p(rays), rays.shape[0], mask, _p(out), nthreads)
        return out

    def trace_any(self, rays, mask=0xFF, nthreads=0):
        rays = np.ascontiguousarray(rays, np.float32)
        out = np.zeros(rays.shape[0], np.uint8)
        lib().orc_trace_any(self.h, _p(rays), rays.shape[0], mask, _p(out), nthreads)
        return out

    def __del__(self):
        if getattr(self, "h", None):
            lib().orc_scene_free(self.h)
            self.h = None


def pathtrace(scene, albedo, emissive, light, camera, width, height, spp, max_depth=8, rr_depth=3, seed=42,
              sequence=3 << 32, sample_offset=0, pixel_range=None, nthreads=0):
    """Oracle path tracer over a committed Scene. albedo: (instances, 3); emissive: (instances,);
    light: (position, u, v, emission); camera: dict(origin, right, up, forward). Returns
    (image float32 (height, width, 4), closest-hit rays, shadow rays); pixels outside
    pixel_range = (begin, end) stay zero."""
    mats = np.zeros(len(emissive), PT_MATERIAL_DTYPE)
    mats["albedo"] = np.asarray(albedo, np.float32).reshape(-1, 3)
    mats["emissive"] = np.asarray(emissive, np.uint32)
    P = _pt_params(light, camera, width, height, spp, max_depth, rr_depth, seed, sequence, sample_offset)
    img = np.zeros((height, width, 4), np.float32)
    counts = np.zeros(2, np.uint64)
    b, e = pixel_range if pixel_range is not None else (0, width * height)
    lib().orc_pathtrace(scene.h, _p(mats), ctypes.byref(P), b, e, _p(img), _p(counts), nthreads)
    return img, int(counts[0]), int(counts[1])


def _pt_params(light, camera, width, height, spp, max_depth, rr_depth, seed, sequence, sample_offset):
    P = PtParams()
    for name, val in zip(("light_position", "light_u", "light_v", "light_emission"), light):
        getattr(P, name)[:] = [float(np.float32(x)) for x in val]
    for name in ("origin", "right", "up", "forward"):
        getattr(P, "camera_" + name)[:] = [float(np.float32(x)) for x in camera[name]]
    P.width, P.height, P.spp, P.sample_offset, P.max_depth, P.rr_depth = width, height, spp, sample_offset, max_depth, rr_depth
    P.seed, P.sequence = seed, sequence
    return P


def pathtrace_rays(scene, albedo, emissive, light, camera, width, height, spp, max_depth=8, rr_depth=3, seed=42,
                   sequence=3 << 32, sample_offset=0, pixel_range=None, cap=1 << 22):
    """The rays the oracle's path tracer traces: (closest-hit rays, shadow rays), (n, 8) float32."""
    mats = np.zeros(len(emissive), PT_MATERIAL_DTYPE)
    mats["albedo"] = np.asarray(albedo, np.float32).reshape(-1, 3)
    mats["emissive"] = np.asarray(emissive, np.uint32)
    P = _pt_params(light, camera, width, height, spp, max_depth, rr_depth, seed, sequence, sample_offset)
    c = np.zeros((cap, 8), np.float32)
    s = np.zeros((cap, 8), np.float32)
    nc, ns = ctypes.c_uint64(), ctypes.c_uint64()
    b, e = pixel_range if pixel_range is not None else (0, width * height)
    lib().orc_pathtrace_rays(scene.h, _p(mats), ctypes.byref(P), b, e, _p(c), cap, ctypes.byref(nc), _p(s), cap,
                             ctypes.byref(ns))
    return c[:min(cap, nc.value)], s[:min(cap, ns.value)]


def sincos_2pi(u):
    s, c = ctypes.c_float(), ctypes.c_float()
    lib().orc_sincos_2pi(float(u), ctypes.byref(s), ctypes.byref(c))
    return s.value, c.value


def single_mesh_scene(verts4, tris, nthreads=0):
    m = Mesh(verts4, tris, nthreads)
    return Scene().add(m).commit()


def brute_closest_f64(verts4, tris, rays, nthreads=0):
    verts4 = np.ascontiguousarray(verts4, np.float32)
    tris = np.ascontiguousarray(tris, np.uint32)
    rays = np.ascontiguousarray(rays, np.float32)
    out = np.zeros(rays.shape[0], BF_DTYPE)
    lib().orc_brute_closest_f64(_p(verts4), verts4.strides[0], _p(tris), tris.shape[0], _p(rays),
                                rays.shape[0], _p(out), nthreads)
    return out


def watertight(o, d, v0, v1, v2, tmin=0.0, tmax=1e30):
    a = [np.ascontiguousarray(x, np.float32) for x in (o, d, v0, v1, v2)]
    t, u, v = ctypes.c_float(), ctypes.c_float(), ctypes.c_float()
    h = lib().orc_watertight(*[_p(x) for x in a], tmin, tmax, ctypes.byref(t), ctypes.byref(u), ctypes.byref(v))
    return bool(h), t.value, u.value, v.value


def offset_ray_origin(p, n):
    """offset_ray_origin (src/dsl/rtx/ray.cpp:16-24) of float32 triples; (k, 3) arrays."""
    p = np.ascontiguousarray(p, np.float32).reshape(-1, 3)
    n = np.ascontiguousarray(n, np.float32).reshape(-1, 3)
    out = np.zeros_like(p)
    for i in range(p.shape[0]):
        lib().orc_offset_ray_origin(_p(p[i]), _p(n[i]), _p(out[i]))
    return out
