"""Accel with managed meshes: the Python frontend's Accel (src/py/luisa/accel.py:58-160) over the
ManagedAccel of src/py/managed_accel.cpp:26-160.

Instances are described by their geometry (vertex / triangle buffer ranges, or an AABB range) and
build option; identical descriptions share one BLAS (the reference keys them by an MD5 of the
description; here the description tuple itself is the key), reference-counted, created on first
use, built at the next update() with PREFER_UPDATE and released when the last instance using it is
removed or re-described. update() builds the BLASes whose description is new and then the TLAS
(PREFER_UPDATE); update_instance_buffer() only rewrites the instance table.

Ray tracing from this object runs as batches over host arrays (trace_closest / trace_any /
query_all / query_any); the per-thread DSL builtins of the reference (BuiltinFuncBuilder) have
their counterparts in the device library of run-time shaders (lc_hip_device.h).
"""
from . import AccelOption, BUILD_PREFER_UPDATE, IDENTITY, LCError, host


class ManagedAccel:
    def __init__(self, device, hint=AccelOption.FAST_BUILD, allow_compact=False, allow_update=False):
        self.device = device
        self.accel = device.create_accel(AccelOption(hint, 1 if allow_compact else 0, 1 if allow_update else 0))
        self._blas = {}         # description -> [blas, refcount]
        self._instances = []    # description per instance
        self._pending = {}      # description -> blas to build at the next update

    # -- BLAS sharing (ManagedAccel::set_mesh / remove_mesh, managed_accel.cpp:22-88)
    def _acquire(self, desc):
        entry = self._blas.get(desc)
        if entry is None:
            kind, option = desc[0], AccelOption(*desc[1])
            if kind == "mesh":
                _, _, vb, voff, vsize, vstride, tb, toff, tsize = desc
                blas = self.device.create_mesh(vb.obj, tb.obj, option=option, vertex_stride=vstride,
                                               vertex_offset=voff, vertex_size=vsize, triangle_offset=toff,
                                               triangle_size=tsize)
            else:
                _, _, ab, aoff, asize = desc
                blas = self.device.create_procedural_primitive(ab.obj, option=option, offset=aoff, size=asize)
            entry = self._blas[desc] = [blas, 0]
            self._pending[desc] = blas
        entry[1] += 1
        return entry[0]

    def _release(self, desc):
        entry = self._blas[desc]
        entry[1] -= 1
        if entry[1] == 0:
            del self._blas[desc]
            self._pending.pop(desc, None)

    @staticmethod
    def _mesh_desc(vb, voff, vsize, vstride, tb, toff, tsize, allow_compact, allow_update, hint):
        # buffers are part of the key by identity (the reference hashes their handles)
        return ("mesh", (hint, 1 if allow_compact else 0, 1 if allow_update else 0), _Ref(vb), voff, vsize, vstride,
                _Ref(tb), toff, tsize)

    @staticmethod
    def _proc_desc(ab, aoff, asize, allow_compact, allow_update, hint):
        return ("procedural", (hint, 1 if allow_compact else 0, 1 if allow_update else 0), _Ref(ab), aoff, asize)

    def _place(self, index, desc, transform, visibility_mask, opaque, user_id):
        blas = self._acquire(desc)
        if index is None:
            self._instances.append(desc)
            self.accel.emplace_back(blas, transform, visibility_mask & 0xFF, opaque, user_id)
        else:
            old = self._instances[index]
            self._instances[index] = desc
            self.accel.set(index, blas, transform, visibility_mask & 0xFF, opaque, user_id)
            self._release(old)

    # -- the frontend's Accel surface (accel.py:68-146)
    def add(self, vertex_buffer, triangle_buffer, transform=IDENTITY, allow_compact=True, allow_update=False,
            visibility_mask=-1, opaque=True, user_id=0, hint=AccelOption.FAST_TRACE):
        self.add_buffer_view(vertex_buffer, 0, vertex_buffer.nbytes, vertex_buffer.stride, triangle_buffer, 0,
                             triangle_buffer.nbytes, transform, allow_compact, allow_update, visibility_mask, opaque,
                             user_id, hint)

    def add_buffer_view(self, vertex_buffer, vertex_byteoffset, vertex_bytesize, vertex_stride, triangle_buffer,
                        triangle_byteoffset, triangle_bytesize, transform=IDENTITY, allow_compact=True,
                        allow_update=False, visibility_mask=-1, opaque=True, user_id=0, hint=AccelOption.FAST_TRACE):
        self._check_view(vertex_buffer, vertex_byteoffset, vertex_bytesize, triangle_buffer, triangle_byteoffset,
                         triangle_bytesize)
        desc = self._mesh_desc(vertex_buffer, vertex_byteoffset, vertex_bytesize, vertex_stride, triangle_buffer,
                               triangle_byteoffset, triangle_bytesize, allow_compact, allow_update, hint)
        self._place(None, desc, transform, visibility_mask, opaque, user_id)

    def add_procedural(self, aabb_buffer, aabb_start_index=0, aabb_count=None, transform=IDENTITY,
                       allow_compact=True, allow_update=False, visibility_mask=-1, opaque=True, user_id=0,
                       hint=AccelOption.FAST_TRACE):
        desc = self._aabb_desc(aabb_buffer, aabb_start_index, aabb_count, allow_compact, allow_update, hint)
        self._place(None, desc, transform, visibility_mask, opaque, user_id)

    def set(self, index, vertex_buffer, triangle_buffer, transform=IDENTITY, allow_compact=True, allow_update=False,
            visibility_mask=-1, opaque=True, user_id=0, hint=AccelOption.FAST_TRACE):
        self.set_buffer_view(index, vertex_buffer, 0, vertex_buffer.nbytes, vertex_buffer.stride, triangle_buffer, 0,
                             triangle_buffer.nbytes, transform, allow_compact, allow_update, visibility_mask, opaque,
                             user_id, hint)

    def set_buffer_view(self, index, vertex_buffer, vertex_byteoffset, vertex_bytesize, vertex_stride,
                        triangle_buffer, triangle_byteoffset, triangle_bytesize, transform=IDENTITY,
                        allow_compact=True, allow_update=False, visibility_mask=-1, opaque=True, user_id=0,
                        hint=AccelOption.FAST_TRACE):
        self._check_view(vertex_buffer, vertex_byteoffset, vertex_bytesize, triangle_buffer, triangle_byteoffset,
                         triangle_bytesize)
        desc = self._mesh_desc(vertex_buffer, vertex_byteoffset, vertex_bytesize, vertex_stride, triangle_buffer,
                               triangle_byteoffset, triangle_bytesize, allow_compact, allow_update, hint)
        self._place(index, desc, transform, visibility_mask, opaque, user_id)

    def set_procedural(self, index, aabb_buffer, aabb_start_index=0, aabb_count=None, transform=IDENTITY,
                       allow_compact=True, allow_update=False, visibility_mask=-1, opaque=True, user_id=0,
                       hint=AccelOption.FAST_TRACE):
        desc = self._aabb_desc(aabb_buffer, aabb_start_index, aabb_count, allow_compact, allow_update, hint)
        self._place(index, desc, transform, visibility_mask, opaque, user_id)

    def pop(self):
        self.accel.pop_back()
        self._release(self._instances.pop())

    def __len__(self):
        return len(self._instances)

    def set_transform_on_update(self, index, transform):
        self.accel.set_transform_on_update(index, transform)

    def set_visibility_on_update(self, index, visibility_mask):
        self.accel.set_visibility_on_update(index, visibility_mask & 0xFF)

    def set_user_id(self, index, user_id):
        self.accel.set_user_id_on_update(index, user_id)

    def update(self, stream, sync=False):
        """Builds the BLASes of new descriptions (PREFER_UPDATE), then the TLAS (managed_accel.cpp:119-157)."""
        for blas in self._pending.values():
            stream << blas.build(BUILD_PREFER_UPDATE)
        self._pending = {}
        stream << self.accel.build(BUILD_PREFER_UPDATE)
        if sync:
            stream.synchronize()

    def update_instance_buffer(self, stream, sync=False):
        stream << self.accel.update_instance_buffer()
        if sync:
            stream.synchronize()

    @property
    def blas_count(self):
        """BLASes alive (shared among the instances that describe the same geometry)."""
        return len(self._blas)

    # -- batches of the frontend's per-thread builtins (accel.py:95-175)
    def trace_closest(self, stream, rays, vis_mask=0xFF):
        return host.trace_closest(self.device, stream, self.accel, rays, vis_mask)

    def trace_any(self, stream, rays, vis_mask=0xFF):
        return host.trace_any(self.device, stream, self.accel, rays, vis_mask)

    def query_all(self, stream, rays, vis_mask=0xFF, **callbacks):
        return host.query(self.device, stream, self.accel, rays, vis_mask, terminate_on_first=False, **callbacks)

    def query_any(self, stream, rays, vis_mask=0xFF, **callbacks):
        return host.query(self.device, stream, self.accel, rays, vis_mask, terminate_on_first=True, **callbacks)

    def instance_transform(self, index):
        return self.accel.instance_info(index)[0]

    def visibility_mask(self, index):
        return self.accel.instance_info(index)[1]

    def user_id(self, index):
        return self.accel.instance_info(index)[2]

    @staticmethod
    def _check_view(vb, voff, vsize, tb, toff, tsize):
        # accel.py:103-106
        if voff + vsize > vb.nbytes or toff + tsize > tb.nbytes:
            raise LCError("buffer view outside its buffer")

    @staticmethod
    def _aabb_desc(ab, start, count, allow_compact, allow_update, hint):
        if ab.stride != 24:
            raise LCError("AABB buffers hold 24-byte elements")  # accel.py:77
        count = ab.count - start if count is None else count
        return ManagedAccel._proc_desc(ab, 24 * start, 24 * count, allow_compact, allow_update, hint)


class _Ref:
    """Identity key of a buffer inside a description (keeps the buffer alive while shared)."""
    __slots__ = ("obj",)

    def __init__(self, obj):
        self.obj = obj

    def __hash__(self):
        return id(self.obj)

    def __eq__(self, other):
        return isinstance(other, _Ref) and other.obj is self.obj
