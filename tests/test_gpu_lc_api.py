"""GPU tests through luisa_compute_lib_interface (include/lc_api.h), called the way the reference's
loader calls an out-of-tree backend (src/backends/common/rust_device_common.cpp:455-480): a context
and a "hip" device from the LibInterface, resources from the DeviceInterface table, and work as C
CommandLists (include/luisa/rust/api_types.h:416-485) with a completion callback. The accel built
through the table is traced and checked against the oracle."""
import ctypes as C
import sys
import threading

import numpy as np
import pytest

import lcapi
import luisacompute_amd as lc
from oracle import binding as ob
from parity import compare_closest

pytestmark = pytest.mark.gpu

COMPUTE = 1  # StreamTag::COMPUTE
FORCE_BUILD = 1


@pytest.fixture(scope="module")
def api():
    dll, li = lcapi.lib_interface(lc.LIB_PATH)
    logs = []

    def log(m):
        logs.append((m.level.decode(), m.message.decode()))
        if logs[-1][0] == "E":  # printed at once: an error is followed by an abort
            print(f"[backend E] {logs[-1][1]}", file=sys.stderr, flush=True)
    cb = lcapi.LoggerCallback(log)
    li.set_logger_callback(cb)
    ctx = li.create_context(b"/tmp")
    dev = li.create_device(ctx, b"hip", b'{"index": 0}')
    assert dev.device.value != 2 ** 64 - 1
    yield dll, li, dev, logs, cb
    dev.destroy_device(dev)
    li.destroy_context(ctx)


def _cmd_list(cmds):
    arr = (lcapi.Command * len(cmds))(*cmds)
    return arr, lcapi.CommandList(arr, len(cmds))


def _upload(buf, data):
    c = lcapi.Command(tag=lcapi.CMD_BUFFER_UPLOAD)
    c.u.buffer_upload = lcapi.BufferUpload(lcapi.Handle(buf), 0, data.nbytes, data.ctypes.data)
    return c


def _download(buf, out):
    c = lcapi.Command(tag=lcapi.CMD_BUFFER_DOWNLOAD)
    c.u.buffer_download = lcapi.BufferDownload(lcapi.Handle(buf), 0, out.nbytes, out.ctypes.data)
    return c


def _dispatch(dev, stream, cmds, callback=None):
    arr, cl = _cmd_list(cmds)
    cb = lcapi.DispatchCallback(callback) if callback else lcapi.DispatchCallback()
    dev.dispatch(dev.device, lcapi.Handle(stream), cl, cb, None)
    return cb  # keep alive until the callback ran


def test_device_table_basics(api):
    dll, li, dev, logs, _ = api
    assert dev.compute_warp_size(dev.device) == 64
    p = dev.query(dev.device, b"device_name")
    name = C.string_at(p).decode()
    li.free_string(p)
    assert name
    # unsupported resources answer with the invalid handle (no abort)
    assert dev.create_shader(dev.device, lcapi.Handle(0), None).resource.handle == 2 ** 64 - 1
    assert dev.create_curve(dev.device, None).handle == 2 ** 64 - 1
    assert dev.create_swapchain(dev.device, None, lcapi.Handle(0)).resource.handle == 2 ** 64 - 1


@pytest.mark.parametrize("make,stride", [
    (lambda: lcapi.ir_vector("float32", 3), 16), (lambda: lcapi.ir_vector("float32", 2), 8),
    (lambda: lcapi.ir_vector("int32", 4), 16), (lambda: lcapi.ir_primitive("uint64"), 8),
    (lambda: lcapi.ir_primitive("bool"), 1), (lambda: lcapi.ir_matrix(3), 48), (lambda: lcapi.ir_matrix(2), 16),
    (lambda: lcapi.ir_struct(40, 8), 40), (lambda: lcapi.ir_array(lcapi.ir_primitive("float32"), 5), 20),
    (lambda: lcapi.ir_array(lcapi.ir_vector("float32", 3), 2), 32)])
def test_create_buffer_element_sizes(api, make, stride):
    """Element strides of IR types (src/rust/luisa_compute_ir/src/ir.rs:207-335)."""
    _, _, dev, _, _ = api
    t = make()
    info = dev.create_buffer(dev.device, C.c_void_p(t.ptr), 7, None)
    assert info.resource.handle != 2 ** 64 - 1 and info.resource.native_handle
    assert (info.element_stride, info.total_size_bytes) == (stride, 7 * stride)
    dev.destroy_buffer(dev.device, lcapi.Handle(info.resource.handle))


def test_void_buffer_is_bytes(api):
    _, _, dev, _, _ = api
    t = lcapi.ir_void()
    info = dev.create_buffer(dev.device, C.c_void_p(t.ptr), 1000, None)
    assert (info.element_stride, info.total_size_bytes) == (0, 1000)
    dev.destroy_buffer(dev.device, lcapi.Handle(info.resource.handle))


def test_build_and_trace_through_command_lists(api):
    """Vertex / index uploads, MeshBuild, ProceduralPrimitiveBuild and AccelBuild (two mesh
    instances and one procedural instance) in one CommandList with a callback; rays uploaded and
    hits downloaded by later lists; the trace runs on the handles the table returned."""
    dll, li, dev, logs, _ = api
    d = dev.device
    verts, tris = lc.icosphere(3)
    f3 = lcapi.ir_vector("float32", 3)
    tri_t = lcapi.ir_struct(12, 4)  # Triangle {uint i0, i1, i2} (include/luisa/runtime/rtx/triangle.h:7-11)
    vb = dev.create_buffer(d, C.c_void_p(f3.ptr), len(verts), None)
    ib = dev.create_buffer(d, C.c_void_p(tri_t.ptr), len(tris), None)
    assert vb.element_stride == 16 and ib.element_stride == 12
    aabbs = np.array([[10, 10, 10, 11, 11, 11]], np.float32)
    # the IrType objects must outlive the calls that read them (a temporary's memory is freed
    # before the call runs)
    void_t = lcapi.ir_void()
    ab = dev.create_buffer(d, C.c_void_p(void_t.ptr), aabbs.nbytes, None)
    st = dev.create_stream(d, COMPUTE).handle
    opt = lcapi.AccelOption(0, False, False)
    mesh = dev.create_mesh(d, C.byref(opt)).handle
    prim = dev.create_procedural_primitive(d, C.byref(opt)).handle
    accel = dev.create_accel(d, C.byref(opt)).handle
    aff1 = np.array([1, 0, 0, 4, 0, 1, 0, 0, 0, 0, 1, 0], np.float32)
    mods = (lcapi.AccelModification * 3)()
    for i, (h, a) in enumerate(((mesh, lc.IDENTITY), (mesh, aff1), (prim, lc.IDENTITY))):
        mods[i].index, mods[i].user_id = i, 100 + i
        mods[i].flags = 1 | 2 | 4 | 16 | 32  # primitive, transform, opaque on, visibility, user id
        mods[i].visibility, mods[i].mesh = 0xFF, h
        mods[i].affine[:] = [float(x) for x in a]
    mb = lcapi.Command(tag=lcapi.CMD_MESH_BUILD)
    mb.u.mesh_build = lcapi.MeshBuild(lcapi.Handle(mesh), FORCE_BUILD, lcapi.Handle(vb.resource.handle), 0,
                                      verts.nbytes, 16, lcapi.Handle(ib.resource.handle), 0, tris.nbytes, 12)
    pb = lcapi.Command(tag=lcapi.CMD_PROCEDURAL_BUILD)
    pb.u.procedural_build = lcapi.ProceduralBuild(lcapi.Handle(prim), FORCE_BUILD, lcapi.Handle(ab.resource.handle), 0, 1)
    acb = lcapi.Command(tag=lcapi.CMD_ACCEL_BUILD)
    acb.u.accel_build = lcapi.AccelBuild(lcapi.Handle(accel), FORCE_BUILD, 3, mods, 3, False)
    done = threading.Event()
    keep = _dispatch(dev, st, [_upload(vb.resource.handle, verts), _upload(ib.resource.handle, tris),
                               _upload(ab.resource.handle, aabbs), mb, pb, acb], lambda user: done.set())
    assert done.wait(60), "dispatch callback did not run"
    rays = lc.rays_box_sphere(50000, initseq=80, lo=-3.0, hi=7.0)
    rbuf = dev.create_buffer(d, C.c_void_p(void_t.ptr), rays.nbytes, None)
    hbuf = dev.create_buffer(d, C.c_void_p(void_t.ptr), 24 * len(rays), None)
    keep2 = _dispatch(dev, st, [_upload(rbuf.resource.handle, rays)])
    rc = dll.lc_hip_accel_trace_closest(C.c_uint64(d.value), C.c_uint64(st), C.c_uint64(accel),
                                        C.c_uint64(rbuf.resource.handle), C.c_size_t(0), C.c_uint64(hbuf.resource.handle),
                                        C.c_size_t(0), C.c_uint32(len(rays)), C.c_uint32(0xFF))
    assert rc == 0
    hits = np.zeros(len(rays), lc.HIT_DTYPE)
    keep3 = _dispatch(dev, st, [_download(hbuf.resource.handle, hits)])
    dev.synchronize_stream(d, lcapi.Handle(st))
    om = ob.Mesh(verts, tris)
    sc = ob.Scene().add(om).add(om, aff1).commit()  # plain traces do not see procedural instances
    rep = compare_closest(hits, sc.trace_closest2(lc.rays_as_floats(rays)), rays)
    assert rep["hits"] > 1000 and (hits["inst"] == 1).sum() > 100
    assert not [m for m in logs if m[0] == "E"]
    del keep, keep2, keep3
    for h in (vb, ib, ab, rbuf, hbuf):
        dev.destroy_buffer(d, lcapi.Handle(h.resource.handle))
    dev.destroy_accel(d, lcapi.Handle(accel))
    dev.destroy_mesh(d, lcapi.Handle(mesh))
    dev.destroy_procedural_primitive(d, lcapi.Handle(prim))
    dev.destroy_stream(d, lcapi.Handle(st))


def test_events_and_textures(api):
    """Timeline events through the table, and image commands: upload, image -> buffer copy,
    buffer -> image copy, image copy and download (RGBA32F = PixelFormat 29, storage FLOAT4 = 14)."""
    _, _, dev, logs, _ = api
    d = dev.device
    st = dev.create_stream(d, COMPUTE).handle
    ev = dev.create_event(d).handle
    assert dev.is_event_completed(d, lcapi.Handle(ev), 0)
    assert not dev.is_event_completed(d, lcapi.Handle(ev), 1)
    w, h = 37, 21
    pix = np.random.default_rng(3).random((h, w, 4), dtype=np.float32)
    t0 = dev.create_texture(d, 29, 2, w, h, 1, 1, False, False).handle
    t1 = dev.create_texture(d, 29, 2, w, h, 1, 1, False, False).handle
    assert t0 != 2 ** 64 - 1 and t1 != 2 ** 64 - 1
    void_t = lcapi.ir_void()
    buf = dev.create_buffer(d, C.c_void_p(void_t.ptr), pix.nbytes, None).resource.handle
    up = lcapi.Command(tag=lcapi.CMD_TEXTURE_UPLOAD)
    up.u.texture_upload = lcapi.TextureTransfer(lcapi.Handle(t0), 14, 0, (C.c_uint32 * 3)(w, h, 1), pix.ctypes.data)
    t2b = lcapi.Command(tag=lcapi.CMD_TEXTURE_TO_BUFFER_COPY)
    t2b.u.texture_to_buffer_copy = lcapi.BufferTextureCopy(lcapi.Handle(buf), 0, lcapi.Handle(t0), 14, 0,
                                                           (C.c_uint32 * 3)(w, h, 1))
    b2t = lcapi.Command(tag=lcapi.CMD_BUFFER_TO_TEXTURE_COPY)
    b2t.u.buffer_to_texture_copy = lcapi.BufferTextureCopy(lcapi.Handle(buf), 0, lcapi.Handle(t1), 14, 0,
                                                           (C.c_uint32 * 3)(w, h, 1))
    tc = lcapi.Command(tag=lcapi.CMD_TEXTURE_COPY)
    tc.u.texture_copy = lcapi.TextureCopy(14, lcapi.Handle(t1), lcapi.Handle(t0), (C.c_uint32 * 3)(w, h, 1), 0, 0)
    got = np.zeros_like(pix)
    down = lcapi.Command(tag=lcapi.CMD_TEXTURE_DOWNLOAD)
    down.u.texture_download = lcapi.TextureTransfer(lcapi.Handle(t0), 14, 0, (C.c_uint32 * 3)(w, h, 1), got.ctypes.data)
    keep = _dispatch(dev, st, [up, t2b, b2t, tc, down])
    dev.signal_event(d, lcapi.Handle(ev), lcapi.Handle(st), 1)
    dev.synchronize_event(d, lcapi.Handle(ev), 1)
    assert dev.is_event_completed(d, lcapi.Handle(ev), 1)
    assert not dev.is_event_completed(d, lcapi.Handle(ev), 2)
    assert np.array_equal(got, pix)
    # bindless slots: a buffer and the image with a bilinear / repeat sampler
    arr = dev.create_bindless_array(d, 4).handle
    mods = (lcapi.BindlessModification * 1)()
    mods[0].slot = 1
    mods[0].buffer = lcapi.BindlessBufferSlot(1, lcapi.Handle(buf), 16)
    mods[0].tex2d = lcapi.BindlessTextureSlot(1, lcapi.Handle(t0), lcapi.Sampler(2, 1))
    bu = lcapi.Command(tag=lcapi.CMD_BINDLESS_UPDATE)
    bu.u.bindless_update = lcapi.BindlessUpdate(lcapi.Handle(arr), mods, 1)
    keep2 = _dispatch(dev, st, [bu])
    dev.synchronize_stream(d, lcapi.Handle(st))
    assert not [m for m in logs if m[0] == "E"]
    del keep, keep2
    dev.destroy_bindless_array(d, lcapi.Handle(arr))
    dev.destroy_texture(d, lcapi.Handle(t0))
    dev.destroy_texture(d, lcapi.Handle(t1))
    dev.destroy_buffer(d, lcapi.Handle(buf))
    dev.destroy_event(d, lcapi.Handle(ev))
    dev.destroy_stream(d, lcapi.Handle(st))
