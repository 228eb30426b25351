"""ctypes mirror of include/lc_api.h (the reference's pure-C backend tables,
include/luisa/rust/api_types.h:12-664), used by the tests that drive the library through
luisa_compute_lib_interface the way the reference's loader does
(src/backends/common/rust_device_common.cpp:455-480). test_lc_api.py checks every size and offset
here against the C header."""
import ctypes as C

u8p = C.POINTER(C.c_uint8)
enum = C.c_int32


class Handle(C.Structure):
    _fields_ = [("value", C.c_uint64)]


class AccelModification(C.Structure):
    _fields_ = [("index", C.c_uint32), ("user_id", C.c_uint32), ("flags", C.c_uint32), ("visibility", C.c_uint32),
                ("mesh", C.c_uint64), ("affine", C.c_float * 12)]


class Sampler(C.Structure):
    _fields_ = [("filter", enum), ("address", enum)]


class BufferUpload(C.Structure):
    _fields_ = [("buffer", Handle), ("offset", C.c_size_t), ("size", C.c_size_t), ("data", C.c_void_p)]


class BufferDownload(C.Structure):
    _fields_ = [("buffer", Handle), ("offset", C.c_size_t), ("size", C.c_size_t), ("data", C.c_void_p)]


class BufferCopy(C.Structure):
    _fields_ = [("src", Handle), ("src_offset", C.c_size_t), ("dst", Handle), ("dst_offset", C.c_size_t),
                ("size", C.c_size_t)]


class BufferTextureCopy(C.Structure):
    _fields_ = [("buffer", Handle), ("buffer_offset", C.c_size_t), ("texture", Handle), ("storage", enum),
                ("texture_level", C.c_uint32), ("texture_size", C.c_uint32 * 3)]


class TextureTransfer(C.Structure):
    _fields_ = [("texture", Handle), ("storage", enum), ("level", C.c_uint32), ("size", C.c_uint32 * 3),
                ("data", C.c_void_p)]


class TextureCopy(C.Structure):
    _fields_ = [("storage", enum), ("src", Handle), ("dst", Handle), ("size", C.c_uint32 * 3),
                ("src_level", C.c_uint32), ("dst_level", C.c_uint32)]


class _ArgBuffer(C.Structure):
    _fields_ = [("buffer", Handle), ("offset", C.c_size_t), ("size", C.c_size_t)]


class _ArgTexture(C.Structure):
    _fields_ = [("texture", Handle), ("level", C.c_uint32)]


class _ArgUniform(C.Structure):
    _fields_ = [("data", C.c_void_p), ("size", C.c_size_t)]


class _ArgUnion(C.Union):
    _fields_ = [("buffer", _ArgBuffer), ("texture", _ArgTexture), ("uniform", _ArgUniform),
                ("bindless_array", Handle), ("accel", Handle)]


class Argument(C.Structure):
    _fields_ = [("tag", enum), ("u", _ArgUnion)]


class ShaderDispatch(C.Structure):
    _fields_ = [("shader", Handle), ("dispatch_size", C.c_uint32 * 3), ("args", C.POINTER(Argument)),
                ("args_count", C.c_size_t)]


class MeshBuild(C.Structure):
    _fields_ = [("mesh", Handle), ("request", enum), ("vertex_buffer", Handle), ("vertex_buffer_offset", C.c_size_t),
                ("vertex_buffer_size", C.c_size_t), ("vertex_stride", C.c_size_t), ("index_buffer", Handle),
                ("index_buffer_offset", C.c_size_t), ("index_buffer_size", C.c_size_t), ("index_stride", C.c_size_t)]


class CurveBuild(C.Structure):
    _fields_ = [("curve", Handle), ("request", enum), ("basis", enum), ("cp_count", C.c_size_t),
                ("seg_count", C.c_size_t), ("cp_buffer", Handle), ("cp_buffer_offset", C.c_size_t),
                ("cp_buffer_stride", C.c_size_t), ("seg_buffer", Handle), ("seg_buffer_offset", C.c_size_t)]


class ProceduralBuild(C.Structure):
    _fields_ = [("primitive", Handle), ("request", enum), ("aabb_buffer", Handle), ("aabb_buffer_offset", C.c_size_t),
                ("aabb_count", C.c_size_t)]


class AccelBuild(C.Structure):
    _fields_ = [("accel", Handle), ("request", enum), ("instance_count", C.c_uint32),
                ("modifications", C.POINTER(AccelModification)), ("modifications_count", C.c_size_t),
                ("update_instance_buffer_only", C.c_bool)]


class BindlessBufferSlot(C.Structure):
    _fields_ = [("op", enum), ("buffer", Handle), ("offset", C.c_size_t)]


class BindlessTextureSlot(C.Structure):
    _fields_ = [("op", enum), ("texture", Handle), ("sampler", Sampler)]


class BindlessModification(C.Structure):
    _fields_ = [("slot", C.c_size_t), ("buffer", BindlessBufferSlot), ("tex2d", BindlessTextureSlot),
                ("tex3d", BindlessTextureSlot)]


class BindlessUpdate(C.Structure):
    _fields_ = [("array", Handle), ("modifications", C.POINTER(BindlessModification)),
                ("modifications_count", C.c_size_t)]


(CMD_BUFFER_UPLOAD, CMD_BUFFER_DOWNLOAD, CMD_BUFFER_COPY, CMD_BUFFER_TO_TEXTURE_COPY, CMD_TEXTURE_TO_BUFFER_COPY,
 CMD_TEXTURE_UPLOAD, CMD_TEXTURE_DOWNLOAD, CMD_TEXTURE_COPY, CMD_SHADER_DISPATCH, CMD_MESH_BUILD, CMD_CURVE_BUILD,
 CMD_PROCEDURAL_BUILD, CMD_ACCEL_BUILD, CMD_BINDLESS_UPDATE) = range(14)


class _CmdUnion(C.Union):
    _fields_ = [("buffer_upload", BufferUpload), ("buffer_download", BufferDownload), ("buffer_copy", BufferCopy),
                ("buffer_to_texture_copy", BufferTextureCopy), ("texture_to_buffer_copy", BufferTextureCopy),
                ("texture_upload", TextureTransfer), ("texture_download", TextureTransfer),
                ("texture_copy", TextureCopy), ("shader_dispatch", ShaderDispatch), ("mesh_build", MeshBuild),
                ("curve_build", CurveBuild), ("procedural_build", ProceduralBuild), ("accel_build", AccelBuild),
                ("bindless_update", BindlessUpdate)]


class Command(C.Structure):
    _fields_ = [("tag", enum), ("u", _CmdUnion)]


class CommandList(C.Structure):
    _fields_ = [("commands", C.POINTER(Command)), ("commands_count", C.c_size_t)]


class CreatedResource(C.Structure):
    _fields_ = [("handle", C.c_uint64), ("native_handle", C.c_void_p)]


class CreatedBuffer(C.Structure):
    _fields_ = [("resource", CreatedResource), ("element_stride", C.c_size_t), ("total_size_bytes", C.c_size_t)]


class CreatedShader(C.Structure):
    _fields_ = [("resource", CreatedResource), ("block_size", C.c_uint32 * 3)]


class CreatedSwapchain(C.Structure):
    _fields_ = [("resource", CreatedResource), ("storage", enum)]


class AccelOption(C.Structure):
    _fields_ = [("hint", enum), ("allow_compaction", C.c_bool), ("allow_update", C.c_bool)]


class ShaderOption(C.Structure):
    _fields_ = [("enable_cache", C.c_bool), ("enable_fast_math", C.c_bool), ("enable_debug_info", C.c_bool),
                ("compile_only", C.c_bool), ("time_trace", C.c_bool), ("max_registers", C.c_uint32),
                ("name", C.c_char_p), ("native_include", C.c_char_p)]


class SwapchainOption(C.Structure):
    _fields_ = [("display", C.c_uint64), ("window", C.c_uint64), ("width", C.c_uint32), ("height", C.c_uint32),
                ("wants_hdr", C.c_bool), ("wants_vsync", C.c_bool), ("back_buffer_count", C.c_uint32)]


class PinnedMemoryExt(C.Structure):
    _fields_ = [("data", C.c_void_p), ("pin_host_memory", C.c_void_p), ("allocate_pinned_memory", C.c_void_p)]


class DenoiserExt(C.Structure):
    _fields_ = [("data", C.c_void_p), ("create", C.c_void_p), ("init", C.c_void_p), ("execute", C.c_void_p),
                ("destroy", C.c_void_p)]


DispatchCallback = C.CFUNCTYPE(None, u8p)


class DeviceInterface(C.Structure):
    pass


DeviceInterface._fields_ = [
    ("device", Handle),
    ("destroy_device", C.CFUNCTYPE(None, DeviceInterface)),
    ("create_buffer", C.CFUNCTYPE(CreatedBuffer, Handle, C.c_void_p, C.c_size_t, C.c_void_p)),
    ("destroy_buffer", C.CFUNCTYPE(None, Handle, Handle)),
    ("create_texture", C.CFUNCTYPE(CreatedResource, Handle, enum, C.c_uint32, C.c_uint32, C.c_uint32, C.c_uint32,
                                   C.c_uint32, C.c_bool, C.c_bool)),
    ("native_handle", C.CFUNCTYPE(C.c_void_p, Handle)),
    ("compute_warp_size", C.CFUNCTYPE(C.c_uint32, Handle)),
    ("destroy_texture", C.CFUNCTYPE(None, Handle, Handle)),
    ("create_bindless_array", C.CFUNCTYPE(CreatedResource, Handle, C.c_size_t)),
    ("destroy_bindless_array", C.CFUNCTYPE(None, Handle, Handle)),
    ("create_stream", C.CFUNCTYPE(CreatedResource, Handle, enum)),
    ("destroy_stream", C.CFUNCTYPE(None, Handle, Handle)),
    ("synchronize_stream", C.CFUNCTYPE(None, Handle, Handle)),
    ("dispatch", C.CFUNCTYPE(None, Handle, Handle, CommandList, DispatchCallback, u8p)),
    ("create_swapchain", C.CFUNCTYPE(CreatedSwapchain, Handle, C.POINTER(SwapchainOption), Handle)),
    ("present_display_in_stream", C.CFUNCTYPE(None, Handle, Handle, Handle, Handle)),
    ("destroy_swapchain", C.CFUNCTYPE(None, Handle, Handle)),
    ("create_shader", C.CFUNCTYPE(CreatedShader, Handle, Handle, C.POINTER(ShaderOption))),
    ("destroy_shader", C.CFUNCTYPE(None, Handle, Handle)),
    ("create_event", C.CFUNCTYPE(CreatedResource, Handle)),
    ("destroy_event", C.CFUNCTYPE(None, Handle, Handle)),
    ("signal_event", C.CFUNCTYPE(None, Handle, Handle, Handle, C.c_uint64)),
    ("synchronize_event", C.CFUNCTYPE(None, Handle, Handle, C.c_uint64)),
    ("wait_event", C.CFUNCTYPE(None, Handle, Handle, Handle, C.c_uint64)),
    ("is_event_completed", C.CFUNCTYPE(C.c_bool, Handle, Handle, C.c_uint64)),
    ("create_mesh", C.CFUNCTYPE(CreatedResource, Handle, C.POINTER(AccelOption))),
    ("destroy_mesh", C.CFUNCTYPE(None, Handle, Handle)),
    ("create_curve", C.CFUNCTYPE(CreatedResource, Handle, C.POINTER(AccelOption))),
    ("destroy_curve", C.CFUNCTYPE(None, Handle, Handle)),
    ("create_procedural_primitive", C.CFUNCTYPE(CreatedResource, Handle, C.POINTER(AccelOption))),
    ("destroy_procedural_primitive", C.CFUNCTYPE(None, Handle, Handle)),
    ("create_accel", C.CFUNCTYPE(CreatedResource, Handle, C.POINTER(AccelOption))),
    ("destroy_accel", C.CFUNCTYPE(None, Handle, Handle)),
    # char* returned as a raw pointer: the caller frees it with LibInterface.free_string
    ("query", C.CFUNCTYPE(C.c_void_p, Handle, C.c_char_p)),
    ("pinned_memory_ext", C.CFUNCTYPE(PinnedMemoryExt, Handle)),
    ("denoiser_ext", C.CFUNCTYPE(DenoiserExt, Handle)),
]


class LoggerMessage(C.Structure):
    _fields_ = [("target", C.c_char_p), ("level", C.c_char_p), ("message", C.c_char_p)]


LoggerCallback = C.CFUNCTYPE(None, LoggerMessage)


class LibInterface(C.Structure):
    _fields_ = [("inner", C.c_void_p), ("set_logger_callback", C.CFUNCTYPE(None, LoggerCallback)),
                ("create_context", C.CFUNCTYPE(Handle, C.c_char_p)), ("destroy_context", C.CFUNCTYPE(None, Handle)),
                ("create_device", C.CFUNCTYPE(DeviceInterface, Handle, C.c_char_p, C.c_char_p)),
                ("free_string", C.CFUNCTYPE(None, C.c_void_p))]


def lib_interface(path):
    """Loads the library and calls its luisa_compute_lib_interface, as the reference loader does."""
    dll = C.CDLL(path)
    fn = dll.luisa_compute_lib_interface
    fn.restype = LibInterface
    fn.argtypes = []
    return dll, fn()


# ---- IR element types (CArc<ir::Type>, include/luisa/rust/ir_common.h:24-46, ir.hpp:138-222),
# built in Python memory for create_buffer. Primitive codes: ir.hpp:20-33.
PRIM = {"bool": 0, "int8": 1, "uint8": 2, "int16": 3, "uint16": 4, "int32": 5, "uint32": 6, "int64": 7, "uint64": 8,
        "float16": 9, "float32": 10, "float64": 11}


class _ArcBlock(C.Structure):
    _fields_ = [("ptr", C.c_void_p), ("ref_count", C.c_size_t), ("destructor", C.c_void_p)]


class IrType:
    """Owns the bytes of one ir::Type and its CArc (keep the object alive while the pointer is used)."""

    def __init__(self, tag, body=b""):
        raw = C.c_int32(tag)
        self.buf = C.create_string_buffer(bytes(raw) + b"\0" * 4 + body + b"\0" * max(0, 40 - len(body)), 56)
        self.block = _ArcBlock(C.cast(self.buf, C.c_void_p), 1, None)
        self.arc = C.c_void_p(C.addressof(self.block))  # CArc<Type> = {CArcSharedBlock<Type> *inner}
        self.keep = []

    @property
    def ptr(self):
        return C.addressof(self.arc)


def ir_void():
    return IrType(0)


def ir_primitive(name):
    return IrType(2, bytes(C.c_int32(PRIM[name])))


def ir_vector(name, n):
    # VectorType {VectorElementType {tag 0 = scalar, body at 8: primitive}, length at 16}
    body = bytes(C.c_int32(0)) + b"\0" * 4 + bytes(C.c_int32(PRIM[name])) + b"\0" * 4 + bytes(C.c_uint32(n))
    return IrType(3, body)


def ir_matrix(n):
    body = bytes(C.c_int32(0)) + b"\0" * 4 + bytes(C.c_int32(PRIM["float32"])) + b"\0" * 4 + bytes(C.c_uint32(n))
    return IrType(4, body)


def ir_struct(size, alignment):
    # StructType {fields: CBoxedSlice (ptr, len, destructor) at 8, alignment at 32, size at 40}
    body = b"\0" * 24 + bytes(C.c_size_t(alignment)) + bytes(C.c_size_t(size))
    return IrType(5, body)


def ir_array(elem, n):
    # ArrayType {element: CArc<Type> at 8, length at 16}
    t = IrType(6, bytes(C.c_void_p(C.addressof(elem.block))) + bytes(C.c_size_t(n)))
    t.keep.append(elem)
    return t
