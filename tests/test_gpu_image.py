"""GPU tests of 2-D images in run-time shaders: uploads / downloads and the reference's texel
conversions (cuda_device_resource.h:225-300) for FLOAT4, HALF4, BYTE4, SHORT2 and INT1 storages."""
import numpy as np
import pytest

import luisacompute_amd as lc
import shaders

pytestmark = pytest.mark.gpu
f32 = np.float32


def _unorm_write(v, scale):
    s = np.clip(np.nan_to_num(v.astype(np.float64), nan=0.0), 0.0, 1.0).astype(np.float32)
    # roundf(saturate(x) * scale): the product rounds in float32, then half away from zero
    return np.floor((s * np.float32(scale)).astype(np.float64) + 0.5)


def test_image_storages_and_conversions(device, stream):
    w, h = 97, 61
    rng = np.random.default_rng(60)
    src = rng.uniform(-0.25, 1.25, (h, w, 4)).astype(np.float32)
    src[0, 0] = [0.5 / 255, 1.5 / 255, 2.5 / 255, np.nan]  # exact halves and a NaN
    src[1, :, :] *= 700.0                                    # beyond the unit range: halves round
    imgs = {k: device.create_image(st, w, h) for k, st in
            (("src", lc.PIXEL_FLOAT4), ("h4", lc.PIXEL_HALF4), ("b4", lc.PIXEL_BYTE4), ("u1", lc.PIXEL_INT1),
             ("s2", lc.PIXEL_SHORT2))}
    stream << imgs["src"].copy_from(src)
    back = device.create_buffer(16, w * h)
    sh = device.create_shader(shaders.IMAGES, "image_kernel", block=(16, 4, 1))
    stream << sh(imgs["src"], imgs["h4"], imgs["b4"], imgs["u1"], imgs["s2"], back).dispatch(w, h)
    out = {k: v.empty_host() for k, v in imgs.items()}
    for k, v in imgs.items():
        stream << v.copy_to(out[k])
    bk = np.zeros((h, w, 4), np.float32)
    stream << back.copy_to(bk) << lc.synchronize()
    np.testing.assert_array_equal(out["src"].view(np.uint32), src.view(np.uint32))  # upload / download
    np.testing.assert_array_equal(out["h4"].view(np.uint16), src.astype(np.float16).view(np.uint16))
    np.testing.assert_array_equal(out["b4"], _unorm_write(src, 255.0).astype(np.uint8))
    np.testing.assert_array_equal(out["s2"], _unorm_write(src[..., :2], 65535.0).astype(np.uint16))
    yy, xx = np.mgrid[0:h, 0:w]
    np.testing.assert_array_equal(out["u1"][..., 0], (xx * 7 + yy).astype(np.int32))
    # reads convert back: unorm bytes x * (1/255), missing channels 0, shorts x * (1/65535)
    b = out["b4"].astype(np.float32) * np.float32(1.0 / 255.0)
    np.testing.assert_array_equal(bk[..., 0], b[..., 0])
    np.testing.assert_array_equal(bk[..., 1], b[..., 3])
    np.testing.assert_array_equal(bk[..., 2], (xx * 7 + yy).astype(np.float32))
    s = out["s2"].astype(np.float32) * np.float32(1.0 / 65535.0)
    np.testing.assert_array_equal(bk[..., 3], s[..., 1] + np.float32(0.0))


def test_image_errors(device, stream):
    with pytest.raises(lc.LCError):
        device.create_image(99, 4, 4)
    img = device.create_image(lc.PIXEL_FLOAT1, 4, 4)
    with pytest.raises(lc.LCError):
        img.copy_from(np.zeros(15, np.float32))


def _np_sample(tex, u, v, code):
    """Exact float32 restatement of the device sampling (point / bilinear, four addresses)."""
    h, w = tex.shape[:2]
    filt, addr = code >> 2, code & 3

    def wrap(i, n):
        if addr == 0:
            return np.clip(i, 0, n - 1)
        if addr == 1:
            return np.mod(i, n)
        if addr == 2:
            m = np.mod(i, 2 * n)
            return np.where(m < n, m, 2 * n - 1 - m)
        return np.where((i < 0) | (i >= n), -1, i)

    def texel(x, y):
        ok = (x >= 0) & (y >= 0)
        return np.where(ok[:, None], tex[np.clip(y, 0, h - 1), np.clip(x, 0, w - 1)], f32(0))

    if filt == 0:
        x = wrap(np.floor(u * f32(w)).astype(np.int64), w)
        y = wrap(np.floor(v * f32(h)).astype(np.int64), h)
        return texel(np.where(y < 0, -1, x), y)
    fx, fy = u * f32(w) - f32(0.5), v * f32(h) - f32(0.5)
    x0, y0 = np.floor(fx), np.floor(fy)
    ax, ay = (fx - x0)[:, None], (fy - y0)[:, None]
    i0, j0 = x0.astype(np.int64), y0.astype(np.int64)
    xa, xb, ya, yb = wrap(i0, w), wrap(i0 + 1, w), wrap(j0, h), wrap(j0 + 1, h)
    t00, t10 = texel(np.where(ya < 0, -1, xa), ya), texel(np.where(ya < 0, -1, xb), ya)
    t01, t11 = texel(np.where(yb < 0, -1, xa), yb), texel(np.where(yb < 0, -1, xb), yb)
    bx, by = f32(1) - ax, f32(1) - ay
    top, bot = t00 * bx + t10 * ax, t01 * bx + t11 * ax
    return top * by + bot * ay


def test_bindless_texture_sampling(device, stream):
    import dsl_kernels
    from luisacompute_amd import host
    rng = np.random.default_rng(61)
    w, h = 13, 7
    tex = rng.uniform(0, 1, (h, w, 4)).astype(f32)
    img = device.create_image(lc.PIXEL_FLOAT4, w, h)
    stream << img.copy_from(tex)
    codes = [lc.sampler_code(f, a) for f in range(4) for a in range(4)]
    heap = device.create_bindless_array(len(codes))
    for k, c in enumerate(codes):
        heap.emplace_texture(k, img, c)
    stream << heap.update()
    n = 4096
    uv = rng.uniform(-1.5, 2.5, (n, 2)).astype(f32)
    uv[:64] = ((np.arange(64)[:, None] % np.array([w, h])) + 0.5) / np.array([w, h], f32)  # texel centres
    k = dsl_kernels.sample_kernel.compile(device)
    ub, ob_ = host.upload(device, stream, uv), device.create_buffer(16, n * len(codes))
    stream << k(ub, ob_, heap, len(codes)).dispatch(n)
    out = np.zeros((n, len(codes), 4), f32)
    stream << ob_.copy_to(out) << lc.synchronize()
    for j, c in enumerate(codes):
        np.testing.assert_array_equal(out[:, j], _np_sample(tex, uv[:, 0], uv[:, 1], c), err_msg=f"sampler {c}")
    # properties that do not restate the formula: texel centres return the texel for every filter
    cx, cy = (np.arange(64) % w), (np.arange(64) % h)
    for j in range(len(codes)):
        np.testing.assert_allclose(out[:64, j], tex[cy, cx], rtol=0, atol=2e-7)
    # zero address: outside the unit square reads 0 for point sampling
    far = (uv[:, 0] < -0.1) | (uv[:, 0] > 1.1) | (uv[:, 1] < -0.2) | (uv[:, 1] > 1.2)
    assert (out[far, codes.index(lc.sampler_code(0, lc.ADDRESS_ZERO))] == 0).all()
    # repeat: u and u + 1 sample the same texel (point filter, away from texel edges)
    j = codes.index(lc.sampler_code(0, lc.ADDRESS_REPEAT))
    fr = np.mod(uv[:, 0] * w, 1.0)
    inner = (fr > 0.01) & (fr < 0.99) & (np.mod(uv[:, 1] * h, 1.0) > 0.01) & (np.mod(uv[:, 1] * h, 1.0) < 0.99)
    xs = np.floor(uv[:, 0] * w).astype(int) % w
    ys = np.floor(uv[:, 1] * h).astype(int) % h
    np.testing.assert_array_equal(out[inner, j], tex[ys[inner], xs[inner]])


def test_volume_shader(device, stream):
    """Volumes (Volume<T>): upload / download, float reads of FLOAT2 (missing channels 0),
    unorm byte writes and reads, int writes, 3-D dispatch coordinates."""
    w, h, d = 19, 7, 5
    rng = np.random.default_rng(12)
    src = rng.uniform(-0.2, 1.2, (d, h, w, 2)).astype(np.float32)
    vs = device.create_volume(lc.PIXEL_FLOAT2, w, h, d)
    vb = device.create_volume(lc.PIXEL_BYTE4, w, h, d)
    vi = device.create_volume(lc.PIXEL_INT1, w, h, d)
    back = device.create_buffer(16, w * h * d)
    stream << vs.copy_from(src)
    sh = device.create_shader(shaders.VOLUMES, "volume_kernel", block=(8, 4, 2))
    stream << sh(vs, vb, vi, back).dispatch(w, h, d)
    ob4, oi1, os2 = vb.empty_host(), vi.empty_host(), vs.empty_host()
    bk = np.zeros((d, h, w, 4), np.float32)
    stream << vb.copy_to(ob4) << vi.copy_to(oi1) << vs.copy_to(os2) << back.copy_to(bk) << lc.synchronize()
    np.testing.assert_array_equal(os2.view(np.uint32), src.view(np.uint32))
    rgba = np.concatenate([src, np.full((d, h, w, 1), 0.5, np.float32), np.ones((d, h, w, 1), np.float32)], axis=-1)
    np.testing.assert_array_equal(ob4, _unorm_write(rgba, 255.0).astype(np.uint8))
    zz, yy, xx = np.mgrid[0:d, 0:h, 0:w]
    np.testing.assert_array_equal(oi1[..., 0], (xx + 10 * yy + 100 * zz - 500).astype(np.int32))
    np.testing.assert_array_equal(bk[..., 0], ob4[..., 0].astype(np.float32) * np.float32(1.0 / 255.0))
    np.testing.assert_array_equal(bk[..., 1], np.zeros((d, h, w), np.float32))
    np.testing.assert_array_equal(bk[..., 2], (xx + 10 * yy + 100 * zz - 500).astype(np.float32))
    np.testing.assert_array_equal(bk[..., 3], np.full((d, h, w), d, np.float32))
    bl = device.create_bindless_array(2)
    with pytest.raises(lc.LCError):
        bl.emplace_texture(0, vs, lc.sampler_code(lc.FILTER_POINT, lc.ADDRESS_EDGE))
        stream << bl.update() << lc.synchronize()


def test_dsl_volume_kernel(device, stream):
    """A Python DSL kernel over volumes (3-D dispatch) equals its numpy restatement."""
    import dsl_kernels
    w, h, d = 13, 6, 4
    src = np.random.default_rng(13).uniform(-2.0, 2.0, (d, h, w, 2)).astype(np.float32)
    vs = device.create_volume(lc.PIXEL_FLOAT2, w, h, d)
    vd = device.create_volume(lc.PIXEL_FLOAT4, w, h, d)
    stream << vs.copy_from(src)
    k = dsl_kernels.volume_kernel.compile(device)
    stream << k(vs, vd, 0.5).dispatch(w, h, d)
    out = vd.empty_host()
    stream << vd.copy_to(out) << lc.synchronize()
    zz = np.broadcast_to(np.arange(d, dtype=np.float32)[:, None, None], (d, h, w))
    np.testing.assert_array_equal(out[..., 0], src[..., 0] * np.float32(0.5))
    np.testing.assert_array_equal(out[..., 1], src[..., 1] * np.float32(0.5))
    np.testing.assert_array_equal(out[..., 2], zz)
    np.testing.assert_array_equal(out[..., 3], np.ones((d, h, w), np.float32))
