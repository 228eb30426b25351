"""GPU tests of 2-D images in run-time shaders: uploads / downloads and the reference's texel
conversions (cuda_device_resource.h:225-300) for FLOAT4, HALF4, BYTE4, SHORT2 and INT1 storages."""
import numpy as np
import pytest

import luisacompute_amd as lc
import shaders

pytestmark = pytest.mark.gpu


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
