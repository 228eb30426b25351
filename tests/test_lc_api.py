"""CPU tests of the pure-C backend entry point (include/lc_api.h): its structures have the layout of
the reference's C header (include/luisa/rust/api_types.h, compiled together with ours when the
reference checkout is present), the ctypes mirror used by the GPU tests matches the C header, and
the library exports luisa_compute_lib_interface with a complete LibInterface."""
import ctypes as C
import os
import subprocess

import pytest

import lcapi

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
REF_HEADER = "/root/reference/include/luisa/rust/api_types.h"
OUR_HEADER = os.path.join(ROOT, "include", "lc_api.h")

# (reference type, our type, [(reference member, our member), ...])
PAIRS = [
    ("LCAccelBuildModification", "lcapi_accel_modification",
     [("index", "index"), ("user_id", "user_id"), ("flags", "flags"), ("visibility", "visibility"), ("mesh", "mesh"),
      ("affine", "affine")]),
    ("LCSampler", "lcapi_sampler", [("filter", "filter"), ("address", "address")]),
    ("LCBufferUploadCommand", "lcapi_buffer_upload", [("buffer", "buffer"), ("offset", "offset"), ("size", "size"), ("data", "data")]),
    ("LCBufferDownloadCommand", "lcapi_buffer_download", [("buffer", "buffer"), ("offset", "offset"), ("size", "size"), ("data", "data")]),
    ("LCBufferCopyCommand", "lcapi_buffer_copy",
     [("src", "src"), ("src_offset", "src_offset"), ("dst", "dst"), ("dst_offset", "dst_offset"), ("size", "size")]),
    ("LCBufferToTextureCopyCommand", "lcapi_buffer_texture_copy",
     [("buffer", "buffer"), ("buffer_offset", "buffer_offset"), ("texture", "texture"), ("storage", "storage"),
      ("texture_level", "texture_level"), ("texture_size", "texture_size")]),
    ("LCTextureToBufferCopyCommand", "lcapi_buffer_texture_copy",
     [("buffer", "buffer"), ("buffer_offset", "buffer_offset"), ("texture", "texture"), ("storage", "storage"),
      ("texture_level", "texture_level"), ("texture_size", "texture_size")]),
    ("LCTextureUploadCommand", "lcapi_texture_transfer",
     [("texture", "texture"), ("storage", "storage"), ("level", "level"), ("size", "size"), ("data", "data")]),
    ("LCTextureDownloadCommand", "lcapi_texture_transfer",
     [("texture", "texture"), ("storage", "storage"), ("level", "level"), ("size", "size"), ("data", "data")]),
    ("LCTextureCopyCommand", "lcapi_texture_copy",
     [("storage", "storage"), ("src", "src"), ("dst", "dst"), ("size", "size"), ("src_level", "src_level"),
      ("dst_level", "dst_level")]),
    ("LCArgument", "lcapi_argument",
     [("tag", "tag"), ("buffer", "u.buffer"), ("buffer.offset", "u.buffer.offset"), ("buffer.size", "u.buffer.size"),
      ("texture", "u.texture"), ("texture.level", "u.texture.level"), ("uniform", "u.uniform"),
      ("uniform.size", "u.uniform.size"), ("bindless_array", "u.bindless_array"), ("accel", "u.accel")]),
    ("LCShaderDispatchCommand", "lcapi_shader_dispatch",
     [("shader", "shader"), ("dispatch_size", "dispatch_size"), ("args", "args"), ("args_count", "args_count")]),
    ("LCMeshBuildCommand", "lcapi_mesh_build",
     [("mesh", "mesh"), ("request", "request"), ("vertex_buffer", "vertex_buffer"),
      ("vertex_buffer_offset", "vertex_buffer_offset"), ("vertex_buffer_size", "vertex_buffer_size"),
      ("vertex_stride", "vertex_stride"), ("index_buffer", "index_buffer"), ("index_buffer_offset", "index_buffer_offset"),
      ("index_buffer_size", "index_buffer_size"), ("index_stride", "index_stride")]),
    ("LCCurveBuildCommand", "lcapi_curve_build",
     [("curve", "curve"), ("request", "request"), ("basis", "basis"), ("cp_count", "cp_count"), ("seg_count", "seg_count"),
      ("cp_buffer", "cp_buffer"), ("cp_buffer_offset", "cp_buffer_offset"), ("cp_buffer_stride", "cp_buffer_stride"),
      ("seg_buffer", "seg_buffer"), ("seg_buffer_offset", "seg_buffer_offset")]),
    ("LCProceduralPrimitiveBuildCommand", "lcapi_procedural_build",
     [("handle", "primitive"), ("request", "request"), ("aabb_buffer", "aabb_buffer"),
      ("aabb_buffer_offset", "aabb_buffer_offset"), ("aabb_count", "aabb_count")]),
    ("LCAccelBuildCommand", "lcapi_accel_build",
     [("accel", "accel"), ("request", "request"), ("instance_count", "instance_count"), ("modifications", "modifications"),
      ("modifications_count", "modifications_count"), ("update_instance_buffer_only", "update_instance_buffer_only")]),
    ("LCBindlessArrayUpdateBuffer", "lcapi_bindless_buffer_slot", [("op", "op"), ("handle", "buffer"), ("offset", "offset")]),
    ("LCBindlessArrayUpdateTexture", "lcapi_bindless_texture_slot", [("op", "op"), ("handle", "texture"), ("sampler", "sampler")]),
    ("LCBindlessArrayUpdateModification", "lcapi_bindless_modification",
     [("slot", "slot"), ("buffer", "buffer"), ("tex2d", "tex2d"), ("tex3d", "tex3d")]),
    ("LCBindlessArrayUpdateCommand", "lcapi_bindless_update",
     [("handle", "array"), ("modifications", "modifications"), ("modifications_count", "modifications_count")]),
    ("LCCommand", "lcapi_command",
     [("tag", "tag"), ("buffer_upload", "u.buffer_upload"), ("buffer_download", "u.buffer_download"),
      ("buffer_copy", "u.buffer_copy"), ("buffer_to_texture_copy", "u.buffer_to_texture_copy"),
      ("texture_to_buffer_copy", "u.texture_to_buffer_copy"), ("texture_upload", "u.texture_upload"),
      ("texture_download", "u.texture_download"), ("texture_copy", "u.texture_copy"),
      ("shader_dispatch", "u.shader_dispatch"), ("mesh_build", "u.mesh_build"), ("curve_build", "u.curve_build"),
      ("procedural_primitive_build", "u.procedural_build"), ("accel_build", "u.accel_build"),
      ("bindless_array_update", "u.bindless_update")]),
    ("LCCommandList", "lcapi_command_list", [("commands", "commands"), ("commands_count", "commands_count")]),
    ("LCCreatedResourceInfo", "lcapi_created_resource", [("handle", "handle"), ("native_handle", "native_handle")]),
    ("LCCreatedBufferInfo", "lcapi_created_buffer",
     [("resource", "resource"), ("element_stride", "element_stride"), ("total_size_bytes", "total_size_bytes")]),
    ("LCCreatedShaderInfo", "lcapi_created_shader", [("resource", "resource"), ("block_size", "block_size")]),
    ("LCCreatedSwapchainInfo", "lcapi_created_swapchain", [("resource", "resource"), ("storage", "storage")]),
    ("LCAccelOption", "lcapi_accel_option", [("hint", "hint"), ("allow_compaction", "allow_compaction"), ("allow_update", "allow_update")]),
    ("LCShaderOption", "lcapi_shader_option",
     [(m, m) for m in ("enable_cache", "enable_fast_math", "enable_debug_info", "compile_only", "time_trace",
                       "max_registers", "name", "native_include")]),
    ("LCSwapchainOption", "lcapi_swapchain_option",
     [(m, m) for m in ("display", "window", "width", "height", "wants_hdr", "wants_vsync", "back_buffer_count")]),
    ("LCPinnedMemoryOption", "lcapi_pinned_memory_option", [("write_combined", "write_combined")]),
    ("LCPinnedMemoryExt", "lcapi_pinned_memory_ext",
     [("data", "data"), ("pin_host_memory", "pin_host_memory"), ("allocate_pinned_memory", "allocate_pinned_memory")]),
    ("LCDenoiserExt", "lcapi_denoiser_ext", [(m, m) for m in ("data", "create", "init", "execute", "destroy")]),
    ("LCDeviceInterface", "lcapi_device_interface",
     [(m, m) for m in ("device", "destroy_device", "create_buffer", "destroy_buffer", "create_texture", "native_handle",
                       "compute_warp_size", "destroy_texture", "create_bindless_array", "destroy_bindless_array",
                       "create_stream", "destroy_stream", "synchronize_stream", "dispatch", "create_swapchain",
                       "present_display_in_stream", "destroy_swapchain", "create_shader", "destroy_shader",
                       "create_event", "destroy_event", "signal_event", "synchronize_event", "wait_event",
                       "is_event_completed", "create_mesh", "destroy_mesh", "create_curve", "destroy_curve",
                       "create_procedural_primitive", "destroy_procedural_primitive", "create_accel",
                       "destroy_accel", "query", "pinned_memory_ext", "denoiser_ext")]),
    ("LCLoggerMessage", "lcapi_logger_message", [("target", "target"), ("level", "level"), ("message", "message")]),
    ("LCLibInterface", "lcapi_lib_interface",
     [(m, m) for m in ("inner", "set_logger_callback", "create_context", "destroy_context", "create_device", "free_string")]),
]

HANDLES = ["LCBuffer", "LCContext", "LCDevice", "LCEvent", "LCStream", "LCShader", "LCSwapchain", "LCTexture",
           "LCBindlessArray", "LCAccel", "LCMesh", "LCCurve", "LCProceduralPrimitive", "LCKernelModule"]

ENUMS = [("LC_COMMAND_BUFFER_UPLOAD", "LCAPI_CMD_BUFFER_UPLOAD"), ("LC_COMMAND_TEXTURE_COPY", "LCAPI_CMD_TEXTURE_COPY"),
         ("LC_COMMAND_SHADER_DISPATCH", "LCAPI_CMD_SHADER_DISPATCH"), ("LC_COMMAND_MESH_BUILD", "LCAPI_CMD_MESH_BUILD"),
         ("LC_COMMAND_PROCEDURAL_PRIMITIVE_BUILD", "LCAPI_CMD_PROCEDURAL_BUILD"),
         ("LC_COMMAND_ACCEL_BUILD", "LCAPI_CMD_ACCEL_BUILD"), ("LC_COMMAND_BINDLESS_ARRAY_UPDATE", "LCAPI_CMD_BINDLESS_UPDATE"),
         ("LC_ACCEL_BUILD_REQUEST_PREFER_UPDATE", "LCAPI_BUILD_PREFER_UPDATE"),
         ("LC_ACCEL_BUILD_REQUEST_FORCE_BUILD", "LCAPI_BUILD_FORCE_BUILD"),
         ("LC_ACCEL_USAGE_HINT_FAST_TRACE", "LCAPI_HINT_FAST_TRACE"), ("LC_ACCEL_USAGE_HINT_FAST_BUILD", "LCAPI_HINT_FAST_BUILD")]


def _compile(src, tmp_path, run=False):
    c = tmp_path / "t.c"
    c.write_text(src)
    exe = tmp_path / "t"
    cmd = ["gcc", "-std=c11", "-Wall", "-Werror", "-I", os.path.join(ROOT, "include"), str(c)]
    cmd += ["-o", str(exe)] if run else ["-fsyntax-only"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
    if run:
        return subprocess.run([str(exe)], capture_output=True, text=True, check=True).stdout


@pytest.mark.skipif(not os.path.exists(REF_HEADER), reason="reference checkout not present")
def test_layout_matches_reference_header(tmp_path):
    lines = [f'#include "{REF_HEADER}"', '#include "lc_api.h"', "#include <stddef.h>", "#include <lc_hip.h>"]
    for ref, ours, members in PAIRS:
        lines.append(f"_Static_assert(sizeof({ref}) == sizeof({ours}), \"size {ref}\");")
        lines.append(f"_Static_assert(_Alignof({ref}) == _Alignof({ours}), \"align {ref}\");")
        for rm, om in members:
            lines.append(f"_Static_assert(offsetof({ref}, {rm}) == offsetof({ours}, {om}), \"{ref}.{rm}\");")
            lines.append(f"_Static_assert(sizeof((({ref} *)0)->{rm}) == sizeof((({ours} *)0)->{om}), \"size {ref}.{rm}\");")
    for h in HANDLES:
        lines.append(f"_Static_assert(sizeof({h}) == sizeof(lcapi_handle), \"{h}\");")
    for a, b in ENUMS:
        lines.append(f"_Static_assert((int){a} == (int){b}, \"{a}\");")
    # modification flag bits (compound-literal macros in the reference header: checked at run time)
    lines.append("int main(void) {")
    for ref_flag, ours in (("PRIMITIVE", "LC_HIP_MOD_PRIMITIVE"), ("TRANSFORM", "LC_HIP_MOD_TRANSFORM"),
                           ("OPAQUE_ON", "LC_HIP_MOD_OPAQUE_ON"), ("OPAQUE_OFF", "LC_HIP_MOD_OPAQUE_OFF"),
                           ("VISIBILITY", "LC_HIP_MOD_VISIBILITY"), ("USER_ID", "LC_HIP_MOD_USER_ID")):
        lines.append(f"  if (LCAccelBuildModificationFlags_{ref_flag}.bits != {ours}) return 1;")
    lines.append('  return 0; }')
    _compile("\n".join(lines) + "\n", tmp_path, run=True)


def test_ctypes_mirror_matches_header(tmp_path):
    """The Python mirror (tests/lcapi.py) against the C header: every size and member offset."""
    mirror = {"lcapi_accel_modification": lcapi.AccelModification, "lcapi_buffer_upload": lcapi.BufferUpload,
              "lcapi_buffer_copy": lcapi.BufferCopy, "lcapi_buffer_texture_copy": lcapi.BufferTextureCopy,
              "lcapi_texture_transfer": lcapi.TextureTransfer, "lcapi_texture_copy": lcapi.TextureCopy,
              "lcapi_argument": lcapi.Argument, "lcapi_shader_dispatch": lcapi.ShaderDispatch,
              "lcapi_mesh_build": lcapi.MeshBuild, "lcapi_curve_build": lcapi.CurveBuild,
              "lcapi_procedural_build": lcapi.ProceduralBuild, "lcapi_accel_build": lcapi.AccelBuild,
              "lcapi_bindless_modification": lcapi.BindlessModification, "lcapi_bindless_update": lcapi.BindlessUpdate,
              "lcapi_command": lcapi.Command, "lcapi_command_list": lcapi.CommandList,
              "lcapi_created_buffer": lcapi.CreatedBuffer, "lcapi_created_shader": lcapi.CreatedShader,
              "lcapi_accel_option": lcapi.AccelOption, "lcapi_shader_option": lcapi.ShaderOption,
              "lcapi_swapchain_option": lcapi.SwapchainOption, "lcapi_device_interface": lcapi.DeviceInterface,
              "lcapi_lib_interface": lcapi.LibInterface}
    lines = ['#include <stdio.h>', '#include <stddef.h>', '#include "lc_api.h"', "int main(void) {"]
    for cname, py in mirror.items():
        lines.append(f'  printf("{cname} __sizeof__ %zu\\n", sizeof({cname}));')
        for fname, _ in py._fields_:
            lines.append(f'  printf("{cname} {fname} %zu\\n", offsetof({cname}, {fname}));')
    lines.append("  return 0; }")
    out = _compile("\n".join(lines) + "\n", tmp_path, run=True)
    for line in out.splitlines():
        cname, what, val = line.split()
        py = mirror[cname]
        got = C.sizeof(py) if what == "__sizeof__" else getattr(py, what).offset
        assert got == int(val), (cname, what, got, val)
    # members of the unions
    assert lcapi.Command.u.offset == 8 and C.sizeof(lcapi.Command) == 88
    assert C.sizeof(lcapi.Argument) == 32


def test_library_exports_lib_interface():
    import luisacompute_amd as lc
    assert os.path.basename(lc.LIB_PATH) == "libluisa_compute_backend_hip.so"
    dll, li = lcapi.lib_interface(lc.LIB_PATH)
    for name, _ in lcapi.LibInterface._fields_[1:]:
        assert C.cast(getattr(li, name), C.c_void_p).value, name
    ctx = li.create_context(b"/tmp")
    assert ctx.value not in (0, 2 ** 64 - 1)
    li.destroy_context(ctx)


def test_unknown_device_name_aborts_with_logged_error(tmp_path):
    """create_device with a name other than "hip" is fatal, as the reference's panic_abort!
    (src/rust/luisa_compute_backend_impl/src/lib.rs:226): the message reaches the logger, then the
    process aborts. Run in a child process (no GPU is touched before the name check)."""
    import sys
    import luisacompute_amd as lc
    prog = (f"import sys; sys.path.insert(0, {os.path.join(ROOT, 'tests')!r}); import lcapi\n"
            f"dll, li = lcapi.lib_interface({lc.LIB_PATH!r})\n"
            "def cb(m): print('LOG', m.level.decode(), m.message.decode(), flush=True)\n"
            "keep = lcapi.LoggerCallback(cb); li.set_logger_callback(keep)\n"
            "ctx = li.create_context(b'/tmp'); li.create_device(ctx, b'cuda', None)\n")
    r = subprocess.run([sys.executable, "-c", prog], capture_output=True, text=True, timeout=120)
    assert r.returncode != 0
    assert "LOG E unknown device cuda" in r.stdout
