"""luisacompute_amd -- MI355X (gfx950) ray-tracing backend for the LuisaCompute programming model.

Python host mirror of the reference runtime's ray-tracing path over the C ABI of
libluisa_compute_backend_hip.so (include/lc_hip.h). Names and argument meaning follow the reference:

    reference (C++)                                   here
    Context::create_device("cuda")                   Device() / create_device("hip")
    device.create_stream()                           Device.create_stream()
    device.create_buffer<T>(n)                       Device.create_buffer(stride, n)
    stream << buffer.copy_from(p) << synchronize()   stream << buf.copy_from(a) ; stream.synchronize()
    device.create_mesh(vb, tb, option)               Device.create_mesh(vb, tb, option)
    device.create_accel(option)                      Device.create_accel(option)
    accel.emplace_back(mesh, transform, mask, ...)   Accel.emplace_back(...)
    stream << mesh.build() << accel.build()          same
    accel.intersect(ray) / intersect_any(ray)        Accel.trace_closest / trace_any over ray buffers
(include/luisa/runtime/{context,device,stream}.h, include/luisa/runtime/rtx/{mesh,accel}.h,
src/runtime/rtx/accel.cpp:68-257, src/py/luisa/accel.py:58-228).

The HIP library is required: there is no CPU fallback. Importing works without a GPU; creating a
Device on a machine without one raises.
"""
import ctypes
import os

import numpy as np

_PKG = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_PKG, "libluisa_compute_backend_hip.so")
INVALID = 0xFFFFFFFFFFFFFFFF

RAY_DTYPE = np.dtype([("origin", "<f4", (3,)), ("t_min", "<f4"), ("direction", "<f4", (3,)), ("t_max", "<f4")],
                     align=True)
HIT_DTYPE = np.dtype([("inst", "<u4"), ("prim", "<u4"), ("bary", "<f4", (2,)), ("committed_ray_t", "<f4"),
                      ("_pad", "<u4")])
# CommittedHit (include/luisa/runtime/rtx/hit.h:18-26): the result of a ray query
COMMITTED_HIT_DTYPE = np.dtype([("inst", "<u4"), ("prim", "<u4"), ("bary", "<f4", (2,)), ("hit_type", "<u4"),
                                ("committed_ray_t", "<f4")])
assert RAY_DTYPE.itemsize == 32 and HIT_DTYPE.itemsize == 24 and COMMITTED_HIT_DTYPE.itemsize == 24
# HitType (include/luisa/runtime/rtx/hit.h:7-14)
HIT_MISS, HIT_SURFACE, HIT_PROCEDURAL = 0, 1, 2
# built-in candidate callbacks of the batched ray query (lc_hip_query_policy)
QUERY_SURFACE_COMMIT, QUERY_SURFACE_IGNORE, QUERY_SURFACE_CUTOUT = 0, 1, 2
QUERY_PROCEDURAL_IGNORE, QUERY_PROCEDURAL_SPHERE = 0, 1


class LCError(RuntimeError):
    pass


class AccelOption(ctypes.Structure):
    """AccelOption (include/luisa/runtime/rhi/resource.h:94-126)."""
    _fields_ = [("hint", ctypes.c_uint32), ("allow_compaction", ctypes.c_uint8), ("allow_update", ctypes.c_uint8),
                ("_pad", ctypes.c_uint8 * 2)]
    FAST_TRACE = 0
    FAST_BUILD = 1


# AccelBuildRequest (include/luisa/runtime/rhi/command.h:402-405); build() defaults to
# PREFER_UPDATE like Mesh::build / Accel::build (include/luisa/runtime/rtx/mesh.h:120)
BUILD_PREFER_UPDATE = 0
BUILD_FORCE_BUILD = 1


class TraceSteps(ctypes.Structure):
    """lc_hip_trace_steps: lane accounting of a counting traversal run."""
    _fields_ = [("nodes_visited", ctypes.c_uint64), ("triangles_tested", ctypes.c_uint64),
                ("live_lane_steps", ctypes.c_uint64), ("wave_steps", ctypes.c_uint64),
                ("refill_rounds", ctypes.c_uint64)]


class QueryPolicy(ctypes.Structure):
    """lc_hip_query_policy: built-in surface / procedural candidate callbacks of a ray query."""
    _fields_ = [("surface_mode", ctypes.c_uint32), ("cutout", ctypes.c_float), ("procedural_mode", ctypes.c_uint32),
                ("terminate_on_first", ctypes.c_uint32)]


class ShaderArgument(ctypes.Structure):
    """lc_hip_shader_argument (Argument, include/luisa/runtime/rhi/argument.h:7-48)."""
    _fields_ = [("tag", ctypes.c_uint32), ("_pad", ctypes.c_uint32), ("handle", ctypes.c_uint64),
                ("offset", ctypes.c_size_t), ("size", ctypes.c_size_t)]
    BUFFER, TEXTURE, UNIFORM, BINDLESS_ARRAY, ACCEL = 0, 1, 2, 3, 4


class BindlessModification(ctypes.Structure):
    """lc_hip_bindless_modification (BindlessArrayUpdateCommand::Modification, command.h:621-680)."""
    _fields_ = [("slot", ctypes.c_size_t), ("buffer", ctypes.c_uint64), ("offset_bytes", ctypes.c_size_t),
                ("buffer_op", ctypes.c_uint32), ("texture_op", ctypes.c_uint32), ("texture", ctypes.c_uint64),
                ("sampler", ctypes.c_uint32), ("_pad", ctypes.c_uint32)]
    NONE, EMPLACE, REMOVE = 0, 1, 2


# Sampler (include/luisa/runtime/rhi/sampler.h): code = filter << 2 | address
FILTER_POINT, FILTER_LINEAR_POINT, FILTER_LINEAR_LINEAR, FILTER_ANISOTROPIC = 0, 1, 2, 3
ADDRESS_EDGE, ADDRESS_REPEAT, ADDRESS_MIRROR, ADDRESS_ZERO = 0, 1, 2, 3


def sampler_code(filter, address):
    return (filter << 2) | address


class Modification(ctypes.Structure):
    """AccelBuildCommand::Modification (include/luisa/runtime/rhi/command.h:534-551)."""
    _fields_ = [("index", ctypes.c_uint32), ("user_id", ctypes.c_uint32), ("flags", ctypes.c_uint32),
                ("vis_mask", ctypes.c_uint32), ("affine", ctypes.c_float * 12), ("primitive", ctypes.c_uint64),
                ("_pad", ctypes.c_uint64)]
    flag_primitive = 1
    flag_transform = 2
    flag_opaque_on = 4
    flag_opaque_off = 8
    flag_visibility = 16
    flag_user_id = 32


class RayBatch(ctypes.Structure):
    """lc_hip_ray_batch: one batch of a batched trace launch."""
    _fields_ = [("ray_buffer", ctypes.c_uint64), ("ray_offset", ctypes.c_size_t), ("hit_buffer", ctypes.c_uint64),
                ("hit_offset", ctypes.c_size_t), ("ray_count", ctypes.c_uint32), ("vis_mask", ctypes.c_uint32)]


class BvhStats(ctypes.Structure):
    _fields_ = [("node_count", ctypes.c_uint32), ("leaf_primitive_count", ctypes.c_uint32), ("depth", ctypes.c_uint32),
                ("primitive_count", ctypes.c_uint32), ("bounds", ctypes.c_float * 6), ("node_bytes", ctypes.c_uint64),
                ("primitive_bytes", ctypes.c_uint64)]


class MemoryUsage(ctypes.Structure):
    """lc_hip_memory_usage: device bytes a built primitive holds."""
    _fields_ = [("node_bytes", ctypes.c_uint64), ("primitive_bytes", ctypes.c_uint64),
                ("scratch_bytes", ctypes.c_uint64), ("total_bytes", ctypes.c_uint64)]


class PtInstance(ctypes.Structure):
    """lc_hip_pt_instance: material of one accel instance (geometry comes from its mesh)."""
    _fields_ = [("albedo", ctypes.c_float * 3), ("emissive", ctypes.c_uint32)]


class PtScene(ctypes.Structure):
    """lc_hip_pt_scene: quad light and pinhole camera."""
    _fields_ = [(n, ctypes.c_float * 3) for n in ("light_position", "light_u", "light_v", "light_emission",
                                                   "camera_origin", "camera_right", "camera_up", "camera_forward")]


class PtOptions(ctypes.Structure):
    """lc_hip_pt_options."""
    _fields_ = [("width", ctypes.c_uint32), ("height", ctypes.c_uint32), ("spp", ctypes.c_uint32),
                ("sample_offset", ctypes.c_uint32), ("max_depth", ctypes.c_uint32), ("rr_depth", ctypes.c_uint32),
                ("seed", ctypes.c_uint64), ("sequence", ctypes.c_uint64), ("mode", ctypes.c_uint32),
                ("chunk_spp", ctypes.c_uint32), ("sort", ctypes.c_uint32), ("graph", ctypes.c_uint32),
                ("fetch_chunk", ctypes.c_uint32), ("fetch_segments", ctypes.c_uint32)]


PT_MEGAKERNEL, PT_WAVEFRONT = 0, 1


class ServerOptions(ctypes.Structure):
    """lc_hip_server_options."""
    _fields_ = [("hbm_budget_bytes", ctypes.c_uint64), ("wave_paths", ctypes.c_uint64), ("max_active", ctypes.c_uint32),
                ("sort", ctypes.c_uint32), ("graph", ctypes.c_uint32), ("fetch_chunk", ctypes.c_uint32),
                ("fetch_segments", ctypes.c_uint32), ("_pad", ctypes.c_uint32)]


class RenderRequest(ctypes.Structure):
    """lc_hip_render_request."""
    _fields_ = [("scene", ctypes.c_uint32), ("width", ctypes.c_uint32), ("height", ctypes.c_uint32),
                ("spp", ctypes.c_uint32), ("seed", ctypes.c_uint64), ("sequence", ctypes.c_uint64),
                ("image_buffer", ctypes.c_uint64), ("image_offset", ctypes.c_size_t)]


class RequestStatus(ctypes.Structure):
    """lc_hip_request_status."""
    _fields_ = [("state", ctypes.c_uint32), ("samples_done", ctypes.c_uint32), ("t_submit", ctypes.c_double),
                ("t_admit", ctypes.c_double), ("t_done", ctypes.c_double)]


class ServingStats(ctypes.Structure):
    """lc_hip_serving_stats."""
    _fields_ = [("waves", ctypes.c_uint64), ("rays_closest", ctypes.c_uint64), ("rays_shadow", ctypes.c_uint64),
                ("requests_done", ctypes.c_uint64), ("peak_active", ctypes.c_uint32), ("active", ctypes.c_uint32),
                ("queued", ctypes.c_uint32), ("_pad", ctypes.c_uint32), ("peak_bytes", ctypes.c_uint64)]


REQUEST_QUEUED, REQUEST_ACTIVE, REQUEST_DONE, REQUEST_FAILED = 0, 1, 2, 3
INSTANCE_TRANSFORM, INSTANCE_USER_ID, INSTANCE_VISIBILITY = 0, 1, 2

assert ctypes.sizeof(Modification) == 80 and ctypes.sizeof(AccelOption) == 8
assert ctypes.sizeof(PtInstance) == 16 and ctypes.sizeof(PtScene) == 96 and ctypes.sizeof(PtOptions) == 64

_lib = None

# name -> (restype, argtypes)
_U64, _U32, _I32, _SZ, _P, _F = ctypes.c_uint64, ctypes.c_uint32, ctypes.c_int32, ctypes.c_size_t, ctypes.c_void_p, ctypes.c_float
_SIGNATURES = {
    "lc_hip_last_error": (ctypes.c_char_p, []),
    "lc_hip_device_count": (ctypes.c_int, []),
    "lc_hip_device_create": (_U64, [_I32]),
    "lc_hip_device_destroy": (None, [_U64]),
    "lc_hip_device_warp_size": (_U32, [_U64]),
    "lc_hip_device_native": (ctypes.c_int, [_U64]),
    "lc_hip_device_query": (ctypes.c_int, [_U64, ctypes.c_char_p, ctypes.c_char_p, _SZ]),
    "lc_hip_buffer_create": (_U64, [_U64, _SZ, _SZ, ctypes.POINTER(_P)]),
    "lc_hip_buffer_destroy": (None, [_U64, _U64]),
    "lc_hip_buffer_create_external": (_U64, [_U64, _SZ, _SZ, _P]),
    "lc_hip_buffer_upload": (ctypes.c_int, [_U64, _U64, _U64, _SZ, _SZ, _P]),
    "lc_hip_buffer_download": (ctypes.c_int, [_U64, _U64, _U64, _SZ, _SZ, _P]),
    "lc_hip_buffer_copy": (ctypes.c_int, [_U64, _U64, _U64, _SZ, _U64, _SZ, _SZ]),
    "lc_hip_stream_create": (_U64, [_U64, _U32]),
    "lc_hip_stream_destroy": (None, [_U64, _U64]),
    "lc_hip_stream_synchronize": (ctypes.c_int, [_U64, _U64]),
    "lc_hip_stream_add_callback": (ctypes.c_int, [_U64, _U64, _P, _P]),
    "lc_hip_stream_native": (_P, [_U64, _U64]),
    "lc_hip_set_name": (ctypes.c_int, [_U64, _U32, _U64, ctypes.c_char_p]),
    "lc_hip_stream_capture_begin": (ctypes.c_int, [_U64, _U64]),
    "lc_hip_stream_capture_end": (ctypes.c_int, [_U64, _U64, ctypes.POINTER(_U64)]),
    "lc_hip_graph_launch": (ctypes.c_int, [_U64, _U64, _U64]),
    "lc_hip_graph_destroy": (None, [_U64, _U64]),
    "lc_hip_event_create": (_U64, [_U64]),
    "lc_hip_event_destroy": (None, [_U64, _U64]),
    "lc_hip_event_signal": (ctypes.c_int, [_U64, _U64, _U64, _U64]),
    "lc_hip_event_wait": (ctypes.c_int, [_U64, _U64, _U64, _U64]),
    "lc_hip_event_synchronize": (ctypes.c_int, [_U64, _U64, _U64]),
    "lc_hip_event_is_completed": (ctypes.c_int, [_U64, _U64, _U64]),
    "lc_hip_mesh_create": (_U64, [_U64, ctypes.POINTER(AccelOption)]),
    "lc_hip_mesh_destroy": (None, [_U64, _U64]),
    "lc_hip_mesh_build": (ctypes.c_int, [_U64, _U64, _U64, _U32, _U64, _SZ, _SZ, _SZ, _U64, _SZ, _SZ]),
    "lc_hip_procedural_create": (_U64, [_U64, ctypes.POINTER(AccelOption)]),
    "lc_hip_procedural_destroy": (None, [_U64, _U64]),
    "lc_hip_procedural_build": (ctypes.c_int, [_U64, _U64, _U64, _U32, _U64, _SZ, _SZ]),
    "lc_hip_accel_query": (ctypes.c_int, [_U64, _U64, _U64, _U64, _SZ, _U64, _SZ, _U32, _U32,
                                          ctypes.POINTER(QueryPolicy)]),
    "lc_hip_texture_create": (_U64, [_U64, _U32, _U32, _U32]),
    "lc_hip_volume_create": (_U64, [_U64, _U32, _U32, _U32, _U32]),
    "lc_hip_texture_destroy": (None, [_U64, _U64]),
    "lc_hip_texture_upload": (ctypes.c_int, [_U64, _U64, _U64, _P]),
    "lc_hip_texture_download": (ctypes.c_int, [_U64, _U64, _U64, _P]),
    "lc_hip_texture_buffer_copy": (ctypes.c_int, [_U64, _U64, _U64, _U64, _SZ, ctypes.c_int]),
    "lc_hip_texture_copy": (ctypes.c_int, [_U64, _U64, _U64, _U64]),
    "lc_hip_texture_info": (ctypes.c_int, [_U64, _U64, _P, _P]),
    "lc_hip_bindless_create": (_U64, [_U64, _SZ]),
    "lc_hip_bindless_destroy": (None, [_U64, _U64]),
    "lc_hip_bindless_update": (ctypes.c_int, [_U64, _U64, _U64, ctypes.POINTER(BindlessModification), _SZ]),
    "lc_hip_shader_compile": (ctypes.c_int, [ctypes.c_char_p, ctypes.c_char_p, _SZ, ctypes.POINTER(_SZ)]),
    "lc_hip_shader_cache_key": (ctypes.c_uint64, [ctypes.c_char_p]),
    "lc_hip_shader_cache_stats": (ctypes.c_int, [ctypes.POINTER(ctypes.c_uint64), ctypes.POINTER(ctypes.c_uint64)]),
    "lc_hip_shader_create": (_U64, [_U64, ctypes.c_char_p, ctypes.c_char_p, _U32, _U32, _U32, ctypes.c_char_p, _SZ]),
    "lc_hip_shader_set_argument_usage": (ctypes.c_int, [_U64, _U64, ctypes.POINTER(_U32), _SZ]),
    "lc_hip_shader_argument_usage": (ctypes.c_int, [_U64, _U64, _SZ]),
    "lc_hip_shader_destroy": (None, [_U64, _U64]),
    "lc_hip_shader_dispatch_multiple": (ctypes.c_int, [_U64, _U64, _U64, ctypes.POINTER(ShaderArgument), _SZ, _P, _SZ,
                                                       _P, _SZ]),
    "lc_hip_shader_dispatch_indirect": (ctypes.c_int, [_U64, _U64, _U64, ctypes.POINTER(ShaderArgument), _SZ, _P, _SZ,
                                                       _U64, _U32, _U32]),
    "lc_hip_shader_dispatch": (ctypes.c_int, [_U64, _U64, _U64, ctypes.POINTER(ShaderArgument), _SZ, _P, _SZ, _U32,
                                              _U32, _U32]),
    "lc_hip_accel_create": (_U64, [_U64, ctypes.POINTER(AccelOption)]),
    "lc_hip_accel_destroy": (None, [_U64, _U64]),
    "lc_hip_accel_build": (ctypes.c_int, [_U64, _U64, _U64, _U32, _U32, ctypes.POINTER(Modification), _SZ, ctypes.c_int]),
    "lc_hip_accel_instance_info": (ctypes.c_int, [_U64, _U64, _U32, ctypes.POINTER(_F), ctypes.POINTER(_U32), ctypes.POINTER(_U32)]),
    "lc_hip_accel_instance_read": (ctypes.c_int, [_U64, _U64, _U64, _U32, _U64, _SZ, _U64, _SZ, _U32]),
    "lc_hip_accel_instance_write": (ctypes.c_int, [_U64, _U64, _U64, _U32, _U64, _SZ, _U64, _SZ, _U32]),
    "lc_hip_accel_trace_closest": (ctypes.c_int, [_U64, _U64, _U64, _U64, _SZ, _U64, _SZ, _U32, _U32]),
    "lc_hip_accel_trace_closest_noalloc": (ctypes.c_int, [_U64, _U64, _U64, _U64, _SZ, _U64, _SZ, _U32, _U32]),
    "lc_hip_accel_trace_closest_indirect": (ctypes.c_int, [_U64, _U64, _U64, _U64, _SZ, _U64, _SZ, _U32, _U64, _SZ,
                                                           _U32]),
    "lc_hip_accel_trace_any_indirect": (ctypes.c_int, [_U64, _U64, _U64, _U64, _SZ, _U64, _SZ, _U32, _U64, _SZ,
                                                       _U32]),
    "lc_hip_accel_trace_any": (ctypes.c_int, [_U64, _U64, _U64, _U64, _SZ, _U64, _SZ, _U32, _U32]),
    "lc_hip_radix_sort_pairs_temp_bytes": (ctypes.c_int, [_U32, ctypes.POINTER(ctypes.c_size_t)]),
    "lc_hip_radix_sort_pairs": (ctypes.c_int, [_U64, _U64, _U64, _U64, _U64, _U64, _U64, _U32, ctypes.c_int,
                                               ctypes.c_int, _U64, _SZ]),
    "lc_hip_accel_trace_count": (ctypes.c_int, [_U64, _U64, _U64, _U64, _SZ, _U32, _U32, ctypes.c_int,
                                                ctypes.POINTER(_U64), ctypes.POINTER(_U64)]),
    "lc_hip_accel_trace_steps": (ctypes.c_int, [_U64, _U64, _U64, _U64, _SZ, _U32, _U32, ctypes.c_int,
                                                ctypes.POINTER(TraceSteps)]),
    "lc_hip_mesh_stats": (ctypes.c_int, [_U64, _U64, ctypes.POINTER(BvhStats)]),
    "lc_hip_mesh_validate": (ctypes.c_int, [_U64, _U64, _U64, ctypes.POINTER(_U32)]),
    "lc_hip_mesh_memory": (ctypes.c_int, [_U64, _U64, ctypes.POINTER(MemoryUsage)]),
    "lc_hip_device_build_arena_bytes": (_U64, [_U64]),
    "lc_hip_device_release_build_arena": (ctypes.c_int, [_U64]),
    "lc_hip_accel_stats": (ctypes.c_int, [_U64, _U64, ctypes.POINTER(BvhStats)]),
    "lc_hip_trace_set_config": (ctypes.c_int, [_U64, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int]),
    "lc_hip_trace_get_config": (ctypes.c_int, [_U64, ctypes.c_int] + [ctypes.POINTER(ctypes.c_int)] * 4),
    "lc_hip_trace_set_ray_sort": (ctypes.c_int, [_U64, ctypes.c_int, ctypes.c_int]),
    "lc_hip_trace_get_ray_sort": (ctypes.c_int, [_U64, ctypes.c_int, ctypes.POINTER(ctypes.c_int)]),
    "lc_hip_trace_set_fetch": (ctypes.c_int, [_U64, ctypes.c_int, _U32, _U32]),
    "lc_hip_trace_get_fetch": (ctypes.c_int, [_U64, ctypes.c_int, ctypes.POINTER(_U32), ctypes.POINTER(_U32)]),
    "lc_hip_trace_set_steps": (ctypes.c_int, [_U64, ctypes.c_int, _U32, _U32]),
    "lc_hip_trace_get_steps": (ctypes.c_int, [_U64, ctypes.c_int, ctypes.POINTER(_U32), ctypes.POINTER(_U32)]),
    "lc_hip_build_set_split_small": (ctypes.c_int, [_U64, ctypes.c_int]),
    "lc_hip_build_set_sah": (ctypes.c_int, [_U64, ctypes.c_int]),
    "lc_hip_build_set_width": (ctypes.c_int, [_U64, ctypes.c_int]),
    "lc_hip_build_set_collapse_head": (ctypes.c_int, [_U64, ctypes.c_int]),
    "lc_hip_build_get_width": (ctypes.c_int, [_U64]),
    "lc_hip_timing_enable": (ctypes.c_int, [_U64, ctypes.c_int]),
    "lc_hip_timing_get": (ctypes.c_int, [_U64, ctypes.c_char_p, ctypes.POINTER(ctypes.c_double), ctypes.POINTER(_U64)]),
    "lc_hip_timing_reset": (ctypes.c_int, [_U64]),
    "lc_hip_workload_icosphere_counts": (None, [ctypes.c_int, ctypes.POINTER(_U32), ctypes.POINTER(_U32)]),
    "lc_hip_workload_icosphere": (None, [ctypes.c_int, _P, _P]),
    "lc_hip_workload_rays_orbit_to_ball": (None, [_U64, _U64, _U64, _F, _P]),
    "lc_hip_workload_soup": (None, [_U64, _U64, _U64, _F, _F, _P, _P]),
    "lc_hip_workload_rays_box_sphere": (None, [_U64, _U64, _U64, _U64, _F, _F, _P]),
    "lc_hip_workload_instances": (None, [_U64, _U64, _U32, _U32, _F, _F, _F, _P, _P]),
    "lc_hip_workload_batch_sizes": (None, [_U64, _U64, _U64, _U32, _U32, _U32, _P]),
    "lc_hip_accel_trace_closest_batched": (ctypes.c_int, [_U64, _U64, _U64, _P, _SZ]),
    "lc_hip_accel_trace_any_batched": (ctypes.c_int, [_U64, _U64, _U64, _P, _SZ]),
    "lc_hip_pathtrace": (ctypes.c_int, [_U64, _U64, _U64, ctypes.POINTER(PtInstance), _U32, ctypes.POINTER(PtScene),
                                        ctypes.POINTER(PtOptions), _U64, _SZ, _U64]),
    "lc_hip_server_create": (_U64, [_U64, ctypes.POINTER(ServerOptions)]),
    "lc_hip_server_destroy": (None, [_U64, _U64]),
    "lc_hip_server_add_scene": (ctypes.c_int, [_U64, _U64, _U64, ctypes.POINTER(PtInstance), _U32, ctypes.POINTER(PtScene),
                                               _U32, _U32, ctypes.POINTER(_U32)]),
    "lc_hip_server_submit": (ctypes.c_int, [_U64, _U64, ctypes.POINTER(RenderRequest), ctypes.POINTER(_U64)]),
    "lc_hip_server_query": (ctypes.c_int, [_U64, _U64, _U64, ctypes.POINTER(RequestStatus)]),
    "lc_hip_server_wait": (ctypes.c_int, [_U64, _U64, _U64, ctypes.c_double]),
    "lc_hip_server_drain": (ctypes.c_int, [_U64, _U64, ctypes.c_double]),
    "lc_hip_server_stats": (ctypes.c_int, [_U64, _U64, ctypes.POINTER(ServingStats)]),
}
EXPORTED_SYMBOLS = sorted(_SIGNATURES)


def lib():
    """Loads libluisa_compute_backend_hip.so (raises if it has not been built: no fallback exists)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise LCError(f"{LIB_PATH} is missing: build it with `python -m luisacompute_amd.build`")
        L = ctypes.CDLL(LIB_PATH)
        for name, (res, args) in _SIGNATURES.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _lib = L
    return _lib


def _check(rc, what=""):
    if rc != 0:
        raise LCError(f"{what}: {lib().lc_hip_last_error().decode()} (status {rc})")


def _handle(h, what):
    if h == INVALID:
        raise LCError(f"{what}: {lib().lc_hip_last_error().decode()}")
    return h


def _ptr(a):
    return a.ctypes.data_as(ctypes.c_void_p)


# ------------------------------------------------------------------ affine helpers
def affine_from_matrix(m):
    """Row-major 3x4 affine from a 4x4 matrix given as numpy [row][col] (command.h:557-571 encodes
    a column-major float4x4 the same way: affine[4*r+c] = M[r][c])."""
    m = np.asarray(m, np.float32)
    if m.shape == (12,):
        return m.copy()
    if m.shape == (3, 4):
        return m.reshape(12).copy()
    return m[:3, :4].reshape(12).copy()


IDENTITY = np.array([1, 0, 0, 0, 0, 1, 0, 0, 0, 0, 1, 0], np.float32)


# ------------------------------------------------------------------ commands
class Command:
    def execute(self, stream):
        raise NotImplementedError


class _Upload(Command):
    def __init__(self, buf, array, offset):
        self.buf, self.array, self.offset = buf, np.ascontiguousarray(array), offset

    def execute(self, stream):
        _check(lib().lc_hip_buffer_upload(stream.device.handle, stream.handle, self.buf.handle, self.offset,
                                          self.array.nbytes, _ptr(self.array)), "buffer upload")


class _Download(Command):
    def __init__(self, buf, array, offset):
        if not array.flags.c_contiguous:
            raise LCError("download destination must be contiguous")
        self.buf, self.array, self.offset = buf, array, offset

    def execute(self, stream):
        _check(lib().lc_hip_buffer_download(stream.device.handle, stream.handle, self.buf.handle, self.offset,
                                            self.array.nbytes, _ptr(self.array)), "buffer download")


class _Copy(Command):
    def __init__(self, src, src_off, dst, dst_off, size):
        self.a = (src, src_off, dst, dst_off, size)

    def execute(self, stream):
        src, so, dst, do, size = self.a
        _check(lib().lc_hip_buffer_copy(stream.device.handle, stream.handle, src.handle, so, dst.handle, do, size),
               "buffer copy")


class _MeshBuild(Command):
    def __init__(self, mesh, request):
        self.mesh, self.request = mesh, request

    def execute(self, stream):
        m = self.mesh
        _check(lib().lc_hip_mesh_build(stream.device.handle, stream.handle, m.handle, self.request, m.vb.handle,
                                       m.v_offset, m.v_size, m.v_stride, m.tb.handle, m.t_offset, m.t_size),
               "mesh build")


class _ProceduralBuild(Command):
    def __init__(self, prim, request):
        self.prim, self.request = prim, request

    def execute(self, stream):
        p = self.prim
        _check(lib().lc_hip_procedural_build(stream.device.handle, stream.handle, p.handle, self.request, p.ab.handle,
                                             p.offset, p.size), "procedural build")


class _AccelBuild(Command):
    def __init__(self, accel, request, mods, instance_count, update_only):
        self.accel, self.request, self.mods, self.n, self.update_only = accel, request, mods, instance_count, update_only

    def execute(self, stream):
        arr = (Modification * max(1, len(self.mods)))(*self.mods)
        _check(lib().lc_hip_accel_build(stream.device.handle, stream.handle, self.accel.handle, self.request, self.n,
                                        arr, len(self.mods), 1 if self.update_only else 0), "accel build")


class _Synchronize(Command):
    def execute(self, stream):
        stream.synchronize()


def synchronize():
    return _Synchronize()


# ------------------------------------------------------------------ resources
class Device:
    """DeviceInterface owner (one GPU). index < 0 selects the default device."""

    def __init__(self, index=0):
        self.handle = _handle(lib().lc_hip_device_create(index), "device create")

    def __del__(self):
        if getattr(self, "handle", None) is not None and _lib is not None:
            lib().lc_hip_device_destroy(self.handle)
            self.handle = None

    def warp_size(self):
        return lib().lc_hip_device_warp_size(self.handle)

    def build_arena_bytes(self):
        """Bytes of the build arena kept for compacting BLAS builds."""
        return lib().lc_hip_device_build_arena_bytes(self.handle)

    def release_build_arena(self):
        _check(lib().lc_hip_device_release_build_arena(self.handle), "release build arena")

    def query(self, key):
        out = ctypes.create_string_buffer(256)
        _check(lib().lc_hip_device_query(self.handle, key.encode(), out, 256), f"query {key}")
        return out.value.decode()

    def create_stream(self):
        return Stream(self)

    def create_buffer(self, stride, count):
        return Buffer(self, stride, count)

    def create_buffer_from(self, array, stream=None):
        array = np.ascontiguousarray(array)
        stride = array.strides[0] if array.ndim > 0 else array.itemsize
        buf = Buffer(self, stride, array.shape[0] if array.ndim > 0 else 1)
        s = stream or Stream(self)
        s << buf.copy_from(array)
        return buf

    def create_mesh(self, vertex_buffer, triangle_buffer, option=None, vertex_stride=16, vertex_offset=0,
                    vertex_size=None, triangle_offset=0, triangle_size=None):
        return Mesh(self, vertex_buffer, triangle_buffer, option, vertex_stride, vertex_offset, vertex_size,
                    triangle_offset, triangle_size)

    def create_procedural_primitive(self, aabb_buffer, option=None, offset=0, size=None):
        return ProceduralPrimitive(self, aabb_buffer, option, offset, size)

    def create_accel(self, option=None):
        return Accel(self, option)

    def create_image(self, storage, width, height):
        return Image(self, storage, width, height)

    def create_volume(self, storage, width, height, depth):
        return Volume(self, storage, width, height, depth)

    def create_bindless_array(self, slots):
        return BindlessArray(self, slots)

    def create_shader(self, source, entry, block=(128, 1, 1)):
        """Compiles HIP source against the device library (lc_hip_device.h) with hiprtc."""
        return Shader(self, source, entry, block)

    def create_event(self):
        return Event(self)

    # configuration / measurement of this backend
    def set_trace_config(self, kind, block, lds_stack, steps, blocks_per_cu=0):
        _check(lib().lc_hip_trace_set_config(self.handle, 1 if kind == "any" else 0, block, lds_stack, steps,
                                             blocks_per_cu), "trace config")

    def trace_config(self, kind):
        v = [ctypes.c_int() for _ in range(4)]
        _check(lib().lc_hip_trace_get_config(self.handle, 1 if kind == "any" else 0, *[ctypes.byref(x) for x in v]),
               "trace config")
        return dict(block=v[0].value, lds_stack=v[1].value, steps=v[2].value, blocks_per_cu=v[3].value)

    def set_ray_sort(self, kind, enable):
        """Traverse rays in sorted-queue order (origin Morton code, direction octant)."""
        _check(lib().lc_hip_trace_set_ray_sort(self.handle, 1 if kind == "any" else 0, 1 if enable else 0), "ray sort")

    def ray_sort(self, kind):
        v = ctypes.c_int()
        _check(lib().lc_hip_trace_get_ray_sort(self.handle, 1 if kind == "any" else 0, ctypes.byref(v)), "ray sort")
        return bool(v.value)

    def set_fetch(self, kind, chunk, segments):
        """Queue-slot hand-out of persistent trace launches (kind 0 closest, 1 any)."""
        _check(lib().lc_hip_trace_set_fetch(self.handle, 1 if kind == "any" else 0, chunk, segments), "set_fetch")

    def fetch(self, kind):
        c, s = ctypes.c_uint32(), ctypes.c_uint32()
        _check(lib().lc_hip_trace_get_fetch(self.handle, 1 if kind == "any" else 0, ctypes.byref(c), ctypes.byref(s)),
               "fetch")
        return c.value, s.value

    def set_trace_steps(self, kind, tri_quarters, refill_below=0):
        """Traversal step policy: triangle-step threshold (1..4 quarters of the live lanes) and
        early refill below that many live lanes (0 off)."""
        _check(lib().lc_hip_trace_set_steps(self.handle, 1 if kind == "any" else 0, tri_quarters, refill_below),
               "set_trace_steps")

    def trace_steps(self, kind):
        q, r = ctypes.c_uint32(), ctypes.c_uint32()
        _check(lib().lc_hip_trace_get_steps(self.handle, 1 if kind == "any" else 0, ctypes.byref(q), ctypes.byref(r)),
               "trace_steps")
        return q.value, r.value

    def set_build_split_small(self, enable):
        _check(lib().lc_hip_build_set_split_small(self.handle, 1 if enable else 0), "build config")

    def set_build_width(self, width):
        """Children per wide BVH node (2..8) of later builds."""
        _check(lib().lc_hip_build_set_width(self.handle, int(width)), "build config")

    def build_width(self):
        return lib().lc_hip_build_get_width(self.handle)

    def set_build_collapse_head(self, levels):
        """Wide levels launched over the whole GPU before one workgroup walks the rest (0: automatic)."""
        _check(lib().lc_hip_build_set_collapse_head(self.handle, int(levels)), "build config")

    def set_build_sah(self, enable):
        """SAH refinement of FAST_TRACE builds (on by default)."""
        _check(lib().lc_hip_build_set_sah(self.handle, 1 if enable else 0), "build config")

    def enable_timing(self, on=True):
        _check(lib().lc_hip_timing_enable(self.handle, 1 if on else 0), "timing")

    def timing(self, name):
        ms, n = ctypes.c_double(), ctypes.c_uint64()
        _check(lib().lc_hip_timing_get(self.handle, name.encode(), ctypes.byref(ms), ctypes.byref(n)), "timing")
        return ms.value, n.value

    def reset_timing(self):
        _check(lib().lc_hip_timing_reset(self.handle), "timing")


def create_device(name="hip", index=0):
    if name not in ("hip", "HIP"):
        raise LCError(f"backend '{name}' is not provided by this package")
    return Device(index)


class Stream:
    def __init__(self, device, tag=1):
        self.device = device
        self.handle = _handle(lib().lc_hip_stream_create(device.handle, tag), "stream create")
        self._thunk = None
        self._pending = {}
        self._next_callback = 0

    def __del__(self):
        if getattr(self, "handle", None) is not None and _lib is not None and self.device.handle is not None:
            lib().lc_hip_stream_destroy(self.device.handle, self.handle)
            self.handle = None

    def __lshift__(self, cmd):
        if callable(cmd) and not isinstance(cmd, Command):
            self.add_callback(cmd)
        else:
            cmd.execute(self)
        return self

    def synchronize(self):
        _check(lib().lc_hip_stream_synchronize(self.device.handle, self.handle), "stream synchronize")

    def add_callback(self, fn):
        """Runs fn() on a backend thread once the work enqueued so far has completed
        (CUDAStream::callback, src/backends/cuda/cuda_stream.cpp:142-170); callbacks of one stream
        run in the order they were added."""
        if self._thunk is None:
            # one C thunk per stream, alive as long as the stream; `user` names the Python callable
            def run(user):
                f = self._pending.pop(user, None)
                if f is not None:
                    f()
            self._thunk = ctypes.CFUNCTYPE(None, ctypes.c_void_p)(run)
        self._next_callback += 1
        key = self._next_callback
        self._pending[key] = fn
        rc = lib().lc_hip_stream_add_callback(self.device.handle, self.handle, ctypes.cast(self._thunk, ctypes.c_void_p),
                                              ctypes.c_void_p(key))
        if rc != 0:
            self._pending.pop(key, None)
        _check(rc, "stream callback")

    def native(self):
        return lib().lc_hip_stream_native(self.device.handle, self.handle)

    def capture(self):
        """Records the launches issued on this stream inside the `with` block into a Graph instead
        of running them (lc_hip_stream_capture_begin / _end): `with stream.capture() as g: ...`,
        then `g.graph.launch(stream)`. Run the sequence once before capturing it."""
        return _Capture(self)


class Graph:
    """A captured launch sequence (a HIP graph); launch(stream) replays it."""

    def __init__(self, device, handle):
        self.device, self.handle = device, handle

    def launch(self, stream):
        _check(lib().lc_hip_graph_launch(self.device.handle, stream.handle, self.handle), "graph launch")

    def __del__(self):
        if getattr(self, "handle", None) is not None and _lib is not None and self.device.handle is not None:
            lib().lc_hip_graph_destroy(self.device.handle, self.handle)
            self.handle = None


class _Capture:
    def __init__(self, stream):
        self.stream, self.graph = stream, None

    def __enter__(self):
        _check(lib().lc_hip_stream_capture_begin(self.stream.device.handle, self.stream.handle), "capture begin")
        return self

    def __exit__(self, exc_type, exc, tb):
        h = ctypes.c_uint64(0)
        rc = lib().lc_hip_stream_capture_end(self.stream.device.handle, self.stream.handle, ctypes.byref(h))
        if exc_type is not None:  # the block failed: drop what was recorded, report the block's error
            if rc == 0:
                lib().lc_hip_graph_destroy(self.stream.device.handle, h.value)
            return False
        _check(rc, "capture end")
        self.graph = Graph(self.stream.device, h.value)
        return False


class Event:
    """Timeline event (include/luisa/runtime/event.h): signal/wait/synchronize by value."""

    def __init__(self, device):
        self.device = device
        self.handle = _handle(lib().lc_hip_event_create(device.handle), "event create")

    def __del__(self):
        if getattr(self, "handle", None) is not None and _lib is not None and self.device.handle is not None:
            lib().lc_hip_event_destroy(self.device.handle, self.handle)
            self.handle = None

    def signal(self, stream, value):
        _check(lib().lc_hip_event_signal(self.device.handle, self.handle, stream.handle, value), "event signal")

    def wait(self, stream, value):
        _check(lib().lc_hip_event_wait(self.device.handle, self.handle, stream.handle, value), "event wait")

    def synchronize(self, value):
        _check(lib().lc_hip_event_synchronize(self.device.handle, self.handle, value), "event synchronize")

    def is_completed(self, value):
        return lib().lc_hip_event_is_completed(self.device.handle, self.handle, value) == 1

    # stream commands, as the reference's Event::signal() / wait() (src/runtime/event.cpp:7-15)
    def signaled(self, value):
        return _EventCommand(self, value, True)

    def waited(self, value):
        return _EventCommand(self, value, False)


class _EventCommand(Command):
    def __init__(self, event, value, signal):
        self.event, self.value, self.is_signal = event, value, signal

    def execute(self, stream):
        if self.is_signal:
            self.event.signal(stream, self.value)
        else:
            self.event.wait(stream, self.value)


class Buffer:
    def __init__(self, device, stride, count):
        self.device, self.stride, self.count = device, int(stride), int(count)
        self.nbytes = self.stride * self.count
        p = ctypes.c_void_p()
        self.handle = _handle(lib().lc_hip_buffer_create(device.handle, self.stride, self.count, ctypes.byref(p)),
                              "buffer create")
        self.native = p.value

    def __del__(self):
        if getattr(self, "handle", None) is not None and _lib is not None and self.device.handle is not None:
            lib().lc_hip_buffer_destroy(self.device.handle, self.handle)
            self.handle = None

    def copy_from(self, array, offset=0):
        return _Upload(self, array, offset)

    def copy_to(self, array, offset=0):
        return _Download(self, array, offset)

    def copy_from_buffer(self, src, src_offset=0, dst_offset=0, size=None):
        return _Copy(src, src_offset, self, dst_offset, self.nbytes - dst_offset if size is None else size)


class Mesh:
    """BLAS over a vertex buffer (16-byte stride multiple) and a uint32 triangle buffer."""

    def __init__(self, device, vb, tb, option=None, vertex_stride=16, vertex_offset=0, vertex_size=None,
                 triangle_offset=0, triangle_size=None):
        self.device, self.vb, self.tb = device, vb, tb
        self.v_stride, self.v_offset = vertex_stride, vertex_offset
        self.v_size = vb.nbytes - vertex_offset if vertex_size is None else vertex_size
        self.t_offset = triangle_offset
        self.t_size = tb.nbytes - triangle_offset if triangle_size is None else triangle_size
        opt = option or AccelOption(AccelOption.FAST_TRACE, 1, 0)
        self.handle = _handle(lib().lc_hip_mesh_create(device.handle, ctypes.byref(opt)), "mesh create")

    def __del__(self):
        if getattr(self, "handle", None) is not None and _lib is not None and self.device.handle is not None:
            lib().lc_hip_mesh_destroy(self.device.handle, self.handle)
            self.handle = None

    def build(self, request=BUILD_PREFER_UPDATE):
        """Builds, or refits when the mesh allows updates, was built before over the same buffer
        ranges, and the request is PREFER_UPDATE."""
        return _MeshBuild(self, request)

    def stats(self):
        s = BvhStats()
        _check(lib().lc_hip_mesh_stats(self.device.handle, self.handle, ctypes.byref(s)), "mesh stats")
        return s

    def memory(self):
        """Device bytes held: {node_bytes, primitive_bytes, scratch_bytes, total_bytes}."""
        return _memory(self, "mesh memory")

    VALIDATION_CHECKS = ("child_index", "triangle_range", "leaf_in_slot_box", "child_in_slot_box",
                         "record_once", "primitive_once", "empty_slot")

    def validate(self, stream):
        """Device-side BVH invariant check; returns {check: violation count} (all zero when sound)."""
        e = (ctypes.c_uint32 * 7)()
        _check(lib().lc_hip_mesh_validate(self.device.handle, stream.handle, self.handle, e), "mesh validate")
        return dict(zip(self.VALIDATION_CHECKS, e[:]))


class DeviceRadixSort:
    """lcub DeviceRadixSort (include/luisa/backends/ext/cuda/lcub/device_radix_sort.h:15-27) for
    uint32 keys and values, on the library's own radix kernels."""

    @staticmethod
    def temp_storage_size(num_items):
        b = ctypes.c_size_t()
        _check(lib().lc_hip_radix_sort_pairs_temp_bytes(int(num_items), ctypes.byref(b)), "radix sort temp bytes")
        return b.value

    @staticmethod
    def sort_pairs(stream, temp, keys_in, keys_out, values_in, values_out, num_items, begin_bit=0, end_bit=32,
                   count_buffer=None, count_index=0):
        """Stable sort of num_items (key, value) pairs by key bits [begin_bit, end_bit). values_in None
        sorts the item indices. count_buffer: only the first min(count_buffer[count_index],
        num_items) items are sorted, the count being read on the device when the sort runs."""
        dev = stream.device
        _check(lib().lc_hip_radix_sort_pairs(dev.handle, stream.handle, temp.handle, keys_in.handle, keys_out.handle,
                                             values_in.handle if values_in is not None else 0, values_out.handle,
                                             int(num_items), int(begin_bit), int(end_bit),
                                             count_buffer.handle if count_buffer is not None else 0,
                                             4 * int(count_index)), "radix sort")


def shader_cache_stats():
    """(hits, misses) of the process's shader code-object cache."""
    h, m = ctypes.c_uint64(), ctypes.c_uint64()
    lib().lc_hip_shader_cache_stats(ctypes.byref(h), ctypes.byref(m))
    return h.value, m.value


def shader_cache_key(source):
    return lib().lc_hip_shader_cache_key(source.encode())


def compile_shader_source(source):
    """Compile-only check of a shader source (no device needed); returns the code object size."""
    log = ctypes.create_string_buffer(1 << 16)
    size = ctypes.c_size_t()
    rc = lib().lc_hip_shader_compile(source.encode(), log, len(log), ctypes.byref(size))
    if rc != 0:
        raise LCError(f"shader compile failed:\n{log.value.decode(errors='replace')}")
    return size.value


class _Dispatch(Command):
    def __init__(self, shader, args, size, multiple=None, indirect=None):
        self.shader, self.args, self.size = shader, args, size
        self.multiple, self.indirect = multiple, indirect

    def execute(self, stream):
        argv = (ShaderArgument * max(1, len(self.args)))()
        blob = bytearray()
        for i, a in enumerate(self.args):
            if isinstance(a, tuple) and isinstance(a[0], Buffer):  # (buffer, offset[, size])
                argv[i] = ShaderArgument(ShaderArgument.BUFFER, 0, a[0].handle, a[1], a[2] if len(a) > 2 else 0)
            elif isinstance(a, Buffer):
                argv[i] = ShaderArgument(ShaderArgument.BUFFER, 0, a.handle, 0, 0)
            elif isinstance(a, Accel):
                argv[i] = ShaderArgument(ShaderArgument.ACCEL, 0, a.handle, 0, 0)
            elif isinstance(a, Image):
                argv[i] = ShaderArgument(ShaderArgument.TEXTURE, 0, a.handle, 0, 0)
            elif isinstance(a, BindlessArray):
                argv[i] = ShaderArgument(ShaderArgument.BINDLESS_ARRAY, 0, a.handle, 0, 0)
            else:  # uniform: numpy scalar / array or bytes
                b = a if isinstance(a, (bytes, bytearray)) else np.ascontiguousarray(a).tobytes()
                argv[i] = ShaderArgument(ShaderArgument.UNIFORM, 0, 0, len(blob), len(b))
                blob += b
        ub = (ctypes.c_ubyte * max(1, len(blob))).from_buffer_copy(bytes(blob) or b"\0")
        common = (stream.device.handle, stream.handle, self.shader.handle, argv, len(self.args),
                  ctypes.cast(ub, ctypes.c_void_p), len(blob))
        if self.multiple is not None:
            sizes = np.ascontiguousarray([(tuple(s) + (1, 1))[:3] for s in self.multiple], np.uint32).reshape(-1)
            _check(lib().lc_hip_shader_dispatch_multiple(*common, _ptr(sizes), len(self.multiple)), "shader dispatch")
        elif self.indirect is not None:
            buf, offset, max_dispatch = self.indirect
            _check(lib().lc_hip_shader_dispatch_indirect(*common, buf.handle, offset, max_dispatch), "shader dispatch")
        else:
            x, y, z = (tuple(self.size) + (1, 1))[:3]
            _check(lib().lc_hip_shader_dispatch(*common, x, y, z), "shader dispatch")


# PixelStorage (include/luisa/runtime/rhi/pixel.h:14-35)
PIXEL_BYTE1, PIXEL_BYTE2, PIXEL_BYTE4, PIXEL_SHORT1, PIXEL_SHORT2, PIXEL_SHORT4 = 0, 1, 2, 3, 4, 5
PIXEL_INT1, PIXEL_INT2, PIXEL_INT4, PIXEL_HALF1, PIXEL_HALF2, PIXEL_HALF4 = 6, 7, 8, 9, 10, 11
PIXEL_FLOAT1, PIXEL_FLOAT2, PIXEL_FLOAT4 = 12, 13, 14
_PIXEL_DTYPES = [np.uint8, np.uint16, np.int32, np.float16, np.float32]


class _ImageCopy(Command):
    def __init__(self, image, array, upload):
        self.image, self.array, self.upload = image, array, upload

    def execute(self, stream):
        fn = lib().lc_hip_texture_upload if self.upload else lib().lc_hip_texture_download
        _check(fn(stream.device.handle, stream.handle, self.image.handle, _ptr(self.array)),
               "image upload" if self.upload else "image download")


class Image:
    """2-D image of one level (Image<T>, include/luisa/runtime/image.h) in a PixelStorage; host
    arrays are (height, width, channels) of the storage's channel type."""

    def __init__(self, device, storage, width, height):
        self.device, self.storage, self.width, self.height = device, storage, int(width), int(height)
        self.handle = _handle(lib().lc_hip_texture_create(device.handle, storage, self.width, self.height),
                              "image create")

    def __del__(self):
        if getattr(self, "handle", None) is not None and _lib is not None and self.device.handle is not None:
            lib().lc_hip_texture_destroy(self.device.handle, self.handle)
            self.handle = None

    @property
    def dtype(self):
        return _PIXEL_DTYPES[self.storage // 3]

    @property
    def channels(self):
        return (1, 2, 4)[self.storage % 3]

    def empty_host(self):
        return np.zeros((self.height, self.width, self.channels), self.dtype)

    def copy_from(self, array):
        a = np.ascontiguousarray(array, self.dtype)
        if a.size != self.width * self.height * self.channels:
            raise LCError("image upload size mismatch")
        return _ImageCopy(self, a, True)

    def copy_to(self, array):
        if array.dtype != self.dtype or array.size != self.width * self.height * self.channels or \
                not array.flags.c_contiguous:
            raise LCError("image download array mismatch")
        return _ImageCopy(self, array, False)


class Volume(Image):
    """Volume of one level (Volume<T>, include/luisa/runtime/volume.h) in a PixelStorage; host
    arrays are (depth, height, width, channels) of the storage's channel type."""

    def __init__(self, device, storage, width, height, depth):
        self.device, self.storage = device, storage
        self.width, self.height, self.depth = int(width), int(height), int(depth)
        self.handle = _handle(lib().lc_hip_volume_create(device.handle, storage, self.width, self.height, self.depth),
                              "volume create")

    def empty_host(self):
        return np.zeros((self.depth, self.height, self.width, self.channels), self.dtype)

    def copy_from(self, array):
        a = np.ascontiguousarray(array, self.dtype)
        if a.size != self.depth * self.width * self.height * self.channels:
            raise LCError("volume upload size mismatch")
        return _ImageCopy(self, a, True)

    def copy_to(self, array):
        if array.dtype != self.dtype or array.size != self.depth * self.width * self.height * self.channels or \
                not array.flags.c_contiguous:
            raise LCError("volume download array mismatch")
        return _ImageCopy(self, array, False)


class _BindlessUpdate(Command):
    def __init__(self, array, mods):
        self.array, self.mods = array, mods

    def execute(self, stream):
        arr = (BindlessModification * max(1, len(self.mods)))(*self.mods)
        _check(lib().lc_hip_bindless_update(stream.device.handle, stream.handle, self.array.handle, arr,
                                            len(self.mods)), "bindless update")


class BindlessArray:
    """Bindless array of buffer slots (BindlessArray, include/luisa/runtime/bindless_array.h):
    emplace / remove record modifications, update() is the BindlessArrayUpdateCommand."""

    def __init__(self, device, slots):
        self.device, self.slots = device, int(slots)
        self.handle = _handle(lib().lc_hip_bindless_create(device.handle, self.slots), "bindless array create")
        self._mods, self._bound = {}, {}

    def __del__(self):
        if getattr(self, "handle", None) is not None and _lib is not None and self.device.handle is not None:
            lib().lc_hip_bindless_destroy(self.device.handle, self.handle)
            self.handle = None

    def _mod(self, slot):
        m = self._mods.get(slot)
        if m is None:
            m = self._mods[slot] = BindlessModification(slot, 0, 0, 0, 0, 0, 0, 0)
        return m

    def emplace(self, slot, buffer, offset_bytes=0):
        m = self._mod(slot)
        m.buffer, m.offset_bytes, m.buffer_op = buffer.handle, offset_bytes, BindlessModification.EMPLACE
        self._bound[slot] = buffer  # keeps the buffer alive while bound

    def remove(self, slot):
        m = self._mod(slot)
        m.buffer, m.offset_bytes, m.buffer_op = 0, 0, BindlessModification.REMOVE
        self._bound.pop(slot, None)

    def emplace_texture(self, slot, image, sampler):
        """2-D texture slot: image with a sampler code (sampler_code(filter, address))."""
        m = self._mod(slot)
        m.texture, m.sampler, m.texture_op = image.handle, sampler, BindlessModification.EMPLACE
        self._bound[("tex", slot)] = image

    def remove_texture(self, slot):
        m = self._mod(slot)
        m.texture, m.sampler, m.texture_op = 0, 0, BindlessModification.REMOVE
        self._bound.pop(("tex", slot), None)

    def update(self):
        mods = [self._mods[k] for k in sorted(self._mods)]
        self._mods = {}
        return _BindlessUpdate(self, mods)


class Shader:
    """A kernel compiled at run time (create_shader, device_interface.h:151-160): HIP source with an
    `extern "C" __global__` entry taking one Params struct (arguments in order, 16-B aligned, then
    `uint4 ls_kid`), compiled against the device library lc_hip_device.h."""

    def __init__(self, device, source, entry, block=(128, 1, 1)):
        self.device = device
        log = ctypes.create_string_buffer(1 << 16)
        bx, by, bz = (tuple(block) + (1, 1))[:3]
        h = lib().lc_hip_shader_create(device.handle, source.encode(), entry.encode(), bx, by, bz, log, len(log))
        if h == INVALID:
            raise LCError(f"shader create: {lib().lc_hip_last_error().decode(errors='replace')[:4000]}")
        self.handle = h

    def __del__(self):
        if getattr(self, "handle", None) is not None and _lib is not None and self.device.handle is not None:
            lib().lc_hip_shader_destroy(self.device.handle, self.handle)
            self.handle = None

    def set_argument_usage(self, usages):
        """Records the kernel's per-argument Usage (0 none, 1 read, 2 write, 3 both)."""
        arr = (ctypes.c_uint32 * max(1, len(usages)))(*usages)
        _check(lib().lc_hip_shader_set_argument_usage(self.device.handle, self.handle, arr, len(usages)),
               "shader argument usage")

    def argument_usage(self, index):
        """shader_argument_usage (device_interface.h:159)."""
        r = lib().lc_hip_shader_argument_usage(self.device.handle, self.handle, index)
        if r < 0:
            raise LCError(f"shader argument usage: {lib().lc_hip_last_error().decode()}")
        return r

    def __call__(self, *args):
        """shader(args...).dispatch(nx[, ny, nz]) -> stream command (ShaderDispatchCommand)."""
        shader = self

        class _Bound:
            def dispatch(self, *size):
                return _Dispatch(shader, args, size)

            def dispatch_multiple(self, sizes):
                """One launch per size (kernel ids 0..n-1)."""
                return _Dispatch(shader, args, None, multiple=list(sizes))

            def dispatch_indirect(self, indirect_buffer, offset=0, max_dispatch=0xFFFFFFFF):
                """Sizes from an indirect dispatch buffer written on the device."""
                return _Dispatch(shader, args, None, indirect=(indirect_buffer, offset, max_dispatch))
        return _Bound()


class ProceduralPrimitive:
    """BLAS over a buffer of AABBs (6 floats each: min, max; include/luisa/runtime/rtx/aabb.h:7-10),
    ProceduralPrimitive (include/luisa/runtime/rtx/procedural_primitive.h). Instanced by an Accel
    like a Mesh; reported to ray queries only."""

    def __init__(self, device, aabb_buffer, option=None, offset=0, size=None):
        self.device, self.ab, self.offset = device, aabb_buffer, offset
        self.size = aabb_buffer.nbytes - offset if size is None else size
        opt = option or AccelOption(AccelOption.FAST_TRACE, 1, 0)
        self.handle = _handle(lib().lc_hip_procedural_create(device.handle, ctypes.byref(opt)),
                              "procedural primitive create")

    def __del__(self):
        if getattr(self, "handle", None) is not None and _lib is not None and self.device.handle is not None:
            lib().lc_hip_procedural_destroy(self.device.handle, self.handle)
            self.handle = None

    def build(self, request=BUILD_PREFER_UPDATE):
        return _ProceduralBuild(self, request)

    def stats(self):
        s = BvhStats()
        _check(lib().lc_hip_mesh_stats(self.device.handle, self.handle, ctypes.byref(s)), "procedural stats")
        return s

    def memory(self):
        """Device bytes held: {node_bytes, primitive_bytes, scratch_bytes, total_bytes}."""
        return _memory(self, "procedural memory")


def _memory(primitive, what):
    u = MemoryUsage()
    _check(lib().lc_hip_mesh_memory(primitive.device.handle, primitive.handle, ctypes.byref(u)), what)
    return {name: getattr(u, name) for name, _ in MemoryUsage._fields_}


class Accel:
    """TLAS with the reference's modification protocol (src/runtime/rtx/accel.cpp:68-257)."""

    def __init__(self, device, option=None):
        self.device = device
        opt = option or AccelOption(AccelOption.FAST_TRACE, 1, 0)
        self.handle = _handle(lib().lc_hip_accel_create(device.handle, ctypes.byref(opt)), "accel create")
        self._meshes = []
        self._mods = {}

    def __del__(self):
        if getattr(self, "handle", None) is not None and _lib is not None and self.device.handle is not None:
            lib().lc_hip_accel_destroy(self.device.handle, self.handle)
            self.handle = None

    def __len__(self):
        return len(self._meshes)

    def _mod(self, index):
        m = self._mods.get(index)
        if m is None:
            m = Modification()
            m.index = index
            self._mods[index] = m
        return m

    def emplace_back(self, mesh, transform=IDENTITY, visibility_mask=0xFF, opaque=True, user_id=0):
        index = len(self._meshes)
        self._meshes.append(mesh)
        self.set(index, mesh, transform, visibility_mask, opaque, user_id)
        return index

    def set(self, index, mesh, transform=IDENTITY, visibility_mask=0xFF, opaque=True, user_id=0):
        self._meshes[index] = mesh
        m = self._mod(index)
        m.flags = Modification.flag_primitive | Modification.flag_transform | Modification.flag_visibility | \
            Modification.flag_user_id | (Modification.flag_opaque_on if opaque else Modification.flag_opaque_off)
        m.primitive = mesh.handle
        m.affine[:] = affine_from_matrix(transform).tolist()
        m.vis_mask = visibility_mask & 0xFF
        m.user_id = user_id

    def pop_back(self):
        self._meshes.pop()
        self._mods.pop(len(self._meshes), None)

    def set_transform_on_update(self, index, transform):
        m = self._mod(index)
        m.flags |= Modification.flag_transform
        m.affine[:] = affine_from_matrix(transform).tolist()

    def set_visibility_on_update(self, index, visibility_mask):
        m = self._mod(index)
        m.flags |= Modification.flag_visibility
        m.vis_mask = visibility_mask & 0xFF

    def set_user_id_on_update(self, index, user_id):
        m = self._mod(index)
        m.flags |= Modification.flag_user_id
        m.user_id = user_id

    def _take_mods(self):
        mods = [self._mods[k] for k in sorted(self._mods)]
        self._mods = {}
        return mods

    def build(self, request=BUILD_PREFER_UPDATE):
        if not self._meshes:
            raise LCError("Building acceleration structure without instances.")  # accel.cpp:72-73
        return _AccelBuild(self, request, self._take_mods(), len(self._meshes), False)

    def update_instance_buffer(self):
        return _AccelBuild(self, 0, self._take_mods(), len(self._meshes), True)

    def instance_info(self, index):
        a = (ctypes.c_float * 12)()
        vis, uid = ctypes.c_uint32(), ctypes.c_uint32()
        _check(lib().lc_hip_accel_instance_info(self.device.handle, self.handle, index, a, ctypes.byref(vis),
                                                ctypes.byref(uid)), "instance info")
        return np.array(a[:], np.float32), vis.value, uid.value

    # in-kernel instance ops (RAY_TRACING_[SET_]INSTANCE_*) as device batches over index buffers;
    # field: INSTANCE_TRANSFORM (12 floats per item), INSTANCE_USER_ID, INSTANCE_VISIBILITY (uint32)
    def instance_read(self, stream, field, index_buffer, out_buffer, count=None, index_offset=0, out_offset=0):
        n = index_buffer.count if count is None else count
        _check(lib().lc_hip_accel_instance_read(self.device.handle, stream.handle, self.handle, field,
                                                index_buffer.handle, index_offset, out_buffer.handle, out_offset, n),
               "instance read")

    def instance_write(self, stream, field, index_buffer, in_buffer, count=None, index_offset=0, in_offset=0):
        n = index_buffer.count if count is None else count
        _check(lib().lc_hip_accel_instance_write(self.device.handle, stream.handle, self.handle, field,
                                                 index_buffer.handle, index_offset, in_buffer.handle, in_offset, n),
               "instance write")

    def stats(self):
        s = BvhStats()
        _check(lib().lc_hip_accel_stats(self.device.handle, self.handle, ctypes.byref(s)), "accel stats")
        return s

    # CallOp::RAY_TRACING_TRACE_CLOSEST / _TRACE_ANY over a batch of rays
    def trace_closest(self, stream, ray_buffer, hit_buffer, count=None, vis_mask=0xFF, ray_offset=0, hit_offset=0):
        n = ray_buffer.count if count is None else count
        _check(lib().lc_hip_accel_trace_closest(self.device.handle, stream.handle, self.handle, ray_buffer.handle,
                                                ray_offset, hit_buffer.handle, hit_offset, n, vis_mask), "trace_closest")

    def trace_any(self, stream, ray_buffer, hit_buffer, count=None, vis_mask=0xFF, ray_offset=0, hit_offset=0):
        n = ray_buffer.count if count is None else count
        _check(lib().lc_hip_accel_trace_any(self.device.handle, stream.handle, self.handle, ray_buffer.handle,
                                            ray_offset, hit_buffer.handle, hit_offset, n, vis_mask), "trace_any")

    def trace_closest_indirect(self, stream, ray_buffer, hit_buffer, count_buffer, count_index=0, capacity=None,
                               vis_mask=0xFF, ray_offset=0, hit_offset=0):
        """Closest hits of the first min(count, capacity) rays, count = count_buffer[count_index]
        (uint32) read on the device when the launch runs (a queue a kernel appended to)."""
        n = ray_buffer.count if capacity is None else capacity
        _check(lib().lc_hip_accel_trace_closest_indirect(self.device.handle, stream.handle, self.handle,
                                                         ray_buffer.handle, ray_offset, hit_buffer.handle, hit_offset,
                                                         n, count_buffer.handle, 4 * count_index, vis_mask),
               "trace_closest_indirect")

    def trace_any_indirect(self, stream, ray_buffer, hit_buffer, count_buffer, count_index=0, capacity=None,
                           vis_mask=0xFF, ray_offset=0, hit_offset=0):
        n = ray_buffer.count if capacity is None else capacity
        _check(lib().lc_hip_accel_trace_any_indirect(self.device.handle, stream.handle, self.handle,
                                                     ray_buffer.handle, ray_offset, hit_buffer.handle, hit_offset,
                                                     n, count_buffer.handle, 4 * count_index, vis_mask),
               "trace_any_indirect")

    # RayQueryAll / RayQueryAny (accel.traverse / traverse_any) over a batch with built-in callbacks;
    # hits: COMMITTED_HIT_DTYPE per ray
    def query(self, stream, ray_buffer, hit_buffer, count=None, vis_mask=0xFF, surface=QUERY_SURFACE_COMMIT,
              cutout=0.8, procedural=QUERY_PROCEDURAL_SPHERE, terminate_on_first=False, ray_offset=0, hit_offset=0):
        n = ray_buffer.count if count is None else count
        pol = QueryPolicy(surface, cutout, procedural, 1 if terminate_on_first else 0)
        _check(lib().lc_hip_accel_query(self.device.handle, stream.handle, self.handle, ray_buffer.handle, ray_offset,
                                        hit_buffer.handle, hit_offset, n, vis_mask, ctypes.byref(pol)), "query")

    def trace_batched(self, stream, batches, any_hit=False):
        """batches: list of (ray_buffer, hit_buffer, count[, vis_mask]); one persistent launch per 64."""
        arr = (RayBatch * max(1, len(batches)))()
        for i, b in enumerate(batches):
            rb, hb, n = b[0], b[1], b[2]
            arr[i] = RayBatch(rb.handle, 0, hb.handle, 0, n, b[3] if len(b) > 3 else 0xFF)
        fn = lib().lc_hip_accel_trace_any_batched if any_hit else lib().lc_hip_accel_trace_closest_batched
        _check(fn(self.device.handle, stream.handle, self.handle, ctypes.cast(arr, ctypes.c_void_p), len(batches)),
               "trace_batched")

    def pathtrace(self, stream, instances, scene, options, image_buffer, image_offset=0, counters=None):
        """Renders into image_buffer (float4 per pixel). instances: one PtInstance per accel
        instance; scene: PtScene; options: PtOptions; counters: optional Buffer of 2 uint64."""
        arr = (PtInstance * max(1, len(instances)))(*instances)
        _check(lib().lc_hip_pathtrace(self.device.handle, stream.handle, self.handle, arr, len(instances),
                                      ctypes.byref(scene), ctypes.byref(options), image_buffer.handle, image_offset,
                                      counters.handle if counters is not None else 0), "pathtrace")

    def trace_steps(self, stream, ray_buffer, count=None, vis_mask=0xFF, any_hit=False, ray_offset=0):
        """The counting run's lane accounting (TraceSteps): node visits, triangle tests, live
        lane-iterations, wave iterations and refill rounds."""
        n = ray_buffer.count if count is None else count
        out = TraceSteps()
        _check(lib().lc_hip_accel_trace_steps(self.device.handle, stream.handle, self.handle, ray_buffer.handle,
                                              ray_offset, n, vis_mask, 1 if any_hit else 0, ctypes.byref(out)),
               "trace_steps")
        return out

    def trace_count(self, stream, ray_buffer, count=None, vis_mask=0xFF, any_hit=False, ray_offset=0):
        n = ray_buffer.count if count is None else count
        a, b = ctypes.c_uint64(), ctypes.c_uint64()
        _check(lib().lc_hip_accel_trace_count(self.device.handle, stream.handle, self.handle, ray_buffer.handle,
                                              ray_offset, n, vis_mask, 1 if any_hit else 0, ctypes.byref(a),
                                              ctypes.byref(b)), "trace_count")
        return a.value, b.value


def _set_name_for(tag):
    def set_name(self, name):
        """Resource::set_name (DeviceInterface::set_name): the builds / dispatches of a named
        resource are bracketed by marker ranges "<name>::build" etc. (rocprofv3 --marker-trace)."""
        _check(lib().lc_hip_set_name(self.device.handle, tag, self.handle, str(name).encode()), "set_name")
        return self
    return set_name


# Resource::Tag values (include/luisa/runtime/rhi/resource.h:179-200)
Mesh.set_name = _set_name_for(3)
ProceduralPrimitive.set_name = _set_name_for(5)
Accel.set_name = _set_name_for(7)
Shader.set_name = _set_name_for(10)


# ------------------------------------------------------------------ render serving
class Server:
    """Render-serving loop (lc_hip_server_*): resident scenes, HBM-budget admission, continuous
    batching of all admitted requests of a scene into shared wavefront waves."""

    def __init__(self, device, hbm_budget=1 << 30, wave_paths=32 << 20, max_active=64, sort=0, graph=1,
                 fetch=(128, 8)):
        self.device = device
        o = ServerOptions(hbm_budget, wave_paths, max_active, sort, graph, fetch[0], fetch[1], 0)
        self.handle = _handle(lib().lc_hip_server_create(device.handle, ctypes.byref(o)), "server create")
        self._keep = []

    def close(self):
        if getattr(self, "handle", None):
            lib().lc_hip_server_destroy(self.device.handle, self.handle)
            self.handle = None

    def __del__(self):
        self.close()

    def add_scene(self, accel, instances, scene, max_depth=8, rr_depth=3):
        arr = (PtInstance * max(1, len(instances)))(*instances)
        sid = ctypes.c_uint32()
        _check(lib().lc_hip_server_add_scene(self.device.handle, self.handle, accel.handle, arr, len(instances),
                                             ctypes.byref(scene), max_depth, rr_depth, ctypes.byref(sid)), "add_scene")
        self._keep.append(accel)
        return sid.value

    def submit(self, scene, width, height, spp, image_buffer, seed=42, sequence=0, image_offset=0):
        q = RenderRequest(scene, width, height, spp, seed, sequence, image_buffer.handle, image_offset)
        rid = ctypes.c_uint64()
        _check(lib().lc_hip_server_submit(self.device.handle, self.handle, ctypes.byref(q), ctypes.byref(rid)), "submit")
        return rid.value

    def query(self, request_id):
        st = RequestStatus()
        _check(lib().lc_hip_server_query(self.device.handle, self.handle, request_id, ctypes.byref(st)), "query")
        return st

    def wait(self, request_id, timeout=600.0):
        _check(lib().lc_hip_server_wait(self.device.handle, self.handle, request_id, timeout), "wait")

    def drain(self, timeout=600.0):
        _check(lib().lc_hip_server_drain(self.device.handle, self.handle, timeout), "drain")

    def stats(self):
        st = ServingStats()
        _check(lib().lc_hip_server_stats(self.device.handle, self.handle, ctypes.byref(st)), "stats")
        return st


# ------------------------------------------------------------------ synthetic workloads
def icosphere(level):
    nv, nt = ctypes.c_uint32(), ctypes.c_uint32()
    lib().lc_hip_workload_icosphere_counts(level, ctypes.byref(nv), ctypes.byref(nt))
    v = np.zeros((nv.value, 4), np.float32)
    t = np.zeros((nt.value, 3), np.uint32)
    lib().lc_hip_workload_icosphere(level, _ptr(v), _ptr(t))
    return v, t


def rays_orbit_to_ball(n, initstate=42, initseq=1, orbit=3.0):
    r = np.zeros(n, RAY_DTYPE)
    lib().lc_hip_workload_rays_orbit_to_ball(initstate, initseq, n, orbit, _ptr(r))
    return r


def soup(ntris, initstate=42, initseq=2, extent=100.0, sigma=0.5):
    v = np.zeros((3 * ntris, 4), np.float32)
    t = np.zeros((ntris, 3), np.uint32)
    lib().lc_hip_workload_soup(initstate, initseq, ntris, extent, sigma, _ptr(v), _ptr(t))
    return v, t


def rays_box_sphere(n, initstate=42, initseq=2, skip=0, lo=0.0, hi=100.0):
    r = np.zeros(n, RAY_DTYPE)
    lib().lc_hip_workload_rays_box_sphere(initstate, initseq, skip, n, lo, hi, _ptr(r))
    return r


def instances(count, mesh_count, initstate=42, initseq=4, extent=50.0, scale_lo=0.5, scale_hi=2.0):
    """Instance affines (count, 12) and mesh ids: Shoemake rotation, translation, scale."""
    a = np.zeros((count, 12), np.float32)
    m = np.zeros(count, np.uint32)
    lib().lc_hip_workload_instances(initstate, initseq, count, mesh_count, extent, scale_lo, scale_hi, _ptr(a), _ptr(m))
    return a, m


def batch_sizes(count, lo, hi, initstate=42, initseq=4, skip=0):
    s = np.zeros(count, np.uint32)
    lib().lc_hip_workload_batch_sizes(initstate, initseq, skip, count, lo, hi, _ptr(s))
    return s


def rays_as_floats(rays):
    """(n, 8) float32 view of a RAY_DTYPE array."""
    return np.ascontiguousarray(rays).view(np.float32).reshape(-1, 8)
