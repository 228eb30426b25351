"""Host conveniences over the runtime mirror: numpy arrays in, numpy arrays out."""
import numpy as np

from . import (COMMITTED_HIT_DTYPE, HIT_DTYPE, QUERY_PROCEDURAL_SPHERE, QUERY_SURFACE_COMMIT, RAY_DTYPE, Accel, Device,
               IDENTITY, LCError, synchronize)


def upload(device, stream, array):
    array = np.ascontiguousarray(array)
    stride = array.strides[0] if array.ndim > 0 else array.itemsize
    buf = device.create_buffer(stride, array.shape[0] if array.ndim > 0 else 1)
    if array.nbytes:
        stream << buf.copy_from(array)
    return buf


def mesh_from_arrays(device, stream, verts4, tris, build=True, option=None):
    """verts4: (n, 4) float32 (16-B stride); tris: (m, 3) uint32; option: AccelOption or None."""
    verts4 = np.ascontiguousarray(verts4, np.float32)
    tris = np.ascontiguousarray(tris, np.uint32).reshape(-1, 3)
    if verts4.ndim != 2 or verts4.shape[1] != 4:
        raise LCError("vertices must be (n, 4) float32")
    vb = device.create_buffer(16, max(1, verts4.shape[0]))
    tb = device.create_buffer(12, max(1, tris.shape[0]))
    if verts4.size:
        stream << vb.copy_from(verts4)
    if tris.size:
        stream << tb.copy_from(tris)
    mesh = device.create_mesh(vb, tb, option=option, vertex_size=16 * verts4.shape[0],
                              triangle_size=12 * tris.shape[0])
    if build:
        stream << mesh.build()
    return mesh


def procedural_from_aabbs(device, stream, aabbs, build=True, option=None):
    """aabbs: (n, 6) float32 rows min.xyz, max.xyz; returns the built ProceduralPrimitive."""
    aabbs = np.ascontiguousarray(aabbs, np.float32).reshape(-1, 6)
    ab = device.create_buffer(24, max(1, aabbs.shape[0]))
    if aabbs.size:
        stream << ab.copy_from(aabbs)
    prim = device.create_procedural_primitive(ab, option=option, size=24 * aabbs.shape[0])
    if build:
        stream << prim.build()
    return prim


def single_mesh_accel(device, stream, verts4, tris, option=None):
    mesh = mesh_from_arrays(device, stream, verts4, tris, option=option)
    accel = device.create_accel(option)
    accel.emplace_back(mesh, IDENTITY)
    stream << accel.build()
    accel._keep = [mesh]
    return accel


def as_rays(rays):
    rays = np.ascontiguousarray(rays)
    if rays.dtype == RAY_DTYPE:
        return rays
    return np.ascontiguousarray(rays, np.float32).reshape(-1, 8).view(RAY_DTYPE).reshape(-1)


def trace_closest(device, stream, accel, rays, vis_mask=0xFF):
    rays = as_rays(rays)
    n = rays.shape[0]
    rb = upload(device, stream, rays)
    hb = device.create_buffer(24, max(1, n))
    out = np.zeros(n, HIT_DTYPE)
    if n:
        accel.trace_closest(stream, rb, hb, n, vis_mask)
        stream << hb.copy_to(out) << synchronize()
    return out


def trace_any(device, stream, accel, rays, vis_mask=0xFF):
    rays = as_rays(rays)
    n = rays.shape[0]
    rb = upload(device, stream, rays)
    hb = device.create_buffer(1, max(1, n))
    out = np.zeros(n, np.uint8)
    if n:
        accel.trace_any(stream, rb, hb, n, vis_mask)
        stream << hb.copy_to(out) << synchronize()
    return out


def query(device, stream, accel, rays, vis_mask=0xFF, surface=QUERY_SURFACE_COMMIT, cutout=0.8,
          procedural=QUERY_PROCEDURAL_SPHERE, terminate_on_first=False):
    """Ray query per ray with built-in callbacks; COMMITTED_HIT_DTYPE array."""
    rays = as_rays(rays)
    n = rays.shape[0]
    rb = upload(device, stream, rays)
    hb = device.create_buffer(24, max(1, n))
    out = np.zeros(n, COMMITTED_HIT_DTYPE)
    if n:
        accel.query(stream, rb, hb, n, vis_mask, surface, cutout, procedural, terminate_on_first)
        stream << hb.copy_to(out) << synchronize()
    return out
