"""The colour path on the GPU through the C ABI.

* LoadColorData (system_utility.cpp:70-75): kf_load_color_file /
  kf_upload_color_jpeg decode on the host (markers, Huffman) and the device
  (dequantisation, ISLOW inverse DCT, fancy upsampling, YCbCr -> B, G, R).
  The device image must equal the oracle's decode and PIL's (libjpeg-turbo,
  cv::imread's decoder) byte for byte: the reference's frames in
  tests/golden (also fixed by SHA-256 in color_golden.json) and JPEGs PIL
  encodes here (4:4:4, 4:2:2, 4:2:0, grey, restart intervals, optimised
  tables, quality 1 and 100, RGB-coded, odd and tiny image sizes), on
  contexts of the image's size.  Uploads that need more coefficient blocks
  than the last one grow the buffers.
* Errors: an image of another size, a missing file, a progressive JPEG;
  the colour image is left as it was.
* SetUpPointClouds (threeD_viewer.cpp:20-55): kf_point_cloud after a few
  pipeline frames equals the oracle's record for the context's model maps
  and colour image, bit for bit, to host and to device memory.
"""
import hashlib
import io
import json
import os

import numpy as np
import pytest

from conftest import ROOT, say, small_setup

pytestmark = pytest.mark.gpu

PIL = pytest.importorskip("PIL")
from PIL import Image, ImageFile  # noqa: E402

ImageFile.MAXBLOCK = 1 << 25
GOLDEN = os.path.join(ROOT, "tests", "golden")
FRAMES = ("img0024.jpg", "img0100.jpg", "img0300.jpg")


def pil_bgr(data):
    return np.asarray(Image.open(io.BytesIO(data)).convert("RGB"))[:, :, ::-1]


def encode(img, **kw):
    bio = io.BytesIO()
    img.save(bio, "JPEG", **kw)
    return bio.getvalue()


def synthetic(rng, h, w, mode="RGB"):
    yy, xx = np.mgrid[0:h, 0:w]
    base = np.stack([(xx * 7 + yy * 3) % 256, (xx * 2 + yy * 11) % 256, (xx * yy) % 256], -1)
    a = (base // 2 + rng.integers(0, 40, (h, w, 3))).astype(np.uint8)
    im = Image.fromarray(a, "RGB")
    return im.convert("L") if mode == "L" else im


def test_reference_frames_on_the_device(kf, oracle):
    golden = json.load(open(os.path.join(GOLDEN, "color_golden.json")))["frames"]
    vol, prm, K, H, W = kf.preset(kf.PRESET_REFERENCE)
    with kf.KinfuContext(vol, prm, H, W) as c:
        assert not c.color().any()  # cv::Mat::zeros until loaded (system_utility.cpp:13)
        for name in FRAMES:
            path = os.path.join(GOLDEN, name)
            say(f"[gpu] colour: kf_load_color_file {name}")
            c.load_color_file(path)
            got = c.color()
            assert hashlib.sha256(got.tobytes()).hexdigest() == golden[name]["bgr_sha256"], name
            data = open(path, "rb").read()
            np.testing.assert_array_equal(got, oracle.jpeg_decode(data))
            np.testing.assert_array_equal(got, pil_bgr(data))
            say(f"[gpu] colour: kf_upload_color_jpeg {name}")
            c.upload_color_bgr(np.zeros((H, W, 3), np.uint8))
            c.upload_color_jpeg(data)
            np.testing.assert_array_equal(c.color(), got)


@pytest.mark.parametrize("size", [(480, 640), (23, 37), (17, 5), (9, 3), (2, 7)])
def test_encoded_variants_on_the_device(kf, oracle, size):
    h, w = size
    rng = np.random.default_rng(h * 1000 + w)
    if size == (480, 640):
        vol, prm, K, H, W = kf.preset(kf.PRESET_REFERENCE)
    else:
        vol, prm, K, H, W = small_setup(kf, H=h, W=w)
    variants = []
    # 4:2:0 first, then 4:4:4 (twice the blocks: the buffers grow), then the rest
    for sub in (2, 0, 1):
        for q in (1, 75, 100):
            for extra in ({}, {"optimize": True}, {"restart_marker_blocks": 3}):
                variants.append((f"sub{sub} q{q} {tuple(extra)}",
                                 encode(synthetic(rng, h, w), quality=q, subsampling=sub, **extra)))
    variants.append(("grey", encode(synthetic(rng, h, w, "L"), quality=80)))
    variants.append(("grey restart", encode(synthetic(rng, h, w, "L"), quality=80,
                                            restart_marker_rows=1)))
    variants.append(("rgb coded", encode(synthetic(rng, h, w), quality=90, keep_rgb=True)))
    with kf.KinfuContext(vol, prm, H, W) as c:
        for what, data in variants:
            say(f"[gpu] colour {h}x{w}: {what}")
            c.upload_color_jpeg(data)
            got = c.color()
            want = oracle.jpeg_decode(data)
            assert want is not None, what
            assert (got == want).all(), (what, int((got != want).sum()))
            assert (got == pil_bgr(data)).all(), what


def test_color_errors_leave_the_image(kf):
    vol, prm, K, H, W = kf.preset(kf.PRESET_REFERENCE)
    rng = np.random.default_rng(5)
    with kf.KinfuContext(vol, prm, H, W) as c:
        good = encode(synthetic(rng, H, W), quality=85)
        c.upload_color_jpeg(good)
        before = c.color()
        say("[gpu] colour errors")
        with pytest.raises(kf.KinfuError, match="colour image is 64x48"):
            c.upload_color_jpeg(encode(synthetic(rng, 48, 64), quality=85))
        with pytest.raises(kf.KinfuError, match="progressive"):
            c.upload_color_jpeg(encode(synthetic(rng, H, W), progressive=True))
        with pytest.raises(kf.KinfuError, match="error -5"):
            c.load_color_file(os.path.join(GOLDEN, "img9999.jpg"))
        np.testing.assert_array_equal(c.color(), before)


def test_point_cloud_equals_oracle(kf, oracle):
    vol, prm, K, H, W = kf.preset(kf.PRESET_ROOM)
    n = 4
    frames, _ = kf.stream(kf.preset_scene(kf.PRESET_ROOM), K, H, W, 0, n, threads=16)
    rng = np.random.default_rng(7)
    with kf.KinfuContext(vol, prm, H, W) as c:
        c.stream_upload(frames)
        say("[gpu] point cloud: pipeline, then colour")
        c.pipeline_run(K, 0, n)
        c.upload_color_jpeg(encode(synthetic(rng, H, W), quality=90))
        got = c.point_cloud()
        m = c.model_maps()
        bgr = c.color()
        want = oracle.point_cloud(m["x"], m["y"], m["z"], bgr)
        assert (got.view(np.uint32) == want.view(np.uint32)).all()
        assert 0.3 < (m["z"] != 0).mean() < 1.0
        say("[gpu] point cloud to device memory")
        torch = pytest.importorskip("torch")
        dev = torch.zeros(H * W * 4, dtype=torch.float32, device="cuda")
        c.synchronize()
        rc = kf.lib.kf_point_cloud(c.h, kf._abi.C.cast(dev.data_ptr(), kf._abi._F),
                                   kf.KF_MEM_DEVICE)
        assert rc == 0
        c.synchronize()
        np.testing.assert_array_equal(dev.cpu().numpy().view(np.uint32),
                                      got.reshape(-1).view(np.uint32))
