"""The colour path on the CPU: the oracle's JPEG decode and point cloud, and
the host half of the product's decode.

The reference loads colour with cv::imread (system_utility.cpp:70-75), i.e.
libjpeg-turbo's default decompression behind OpenCV; neither is in the
reference.  The oracle restates that decoder (oracle/kf_oracle_jpeg.cpp) and
is pinned here against PIL, which decodes with libjpeg-turbo and the same
defaults (ISLOW inverse DCT, fancy upsampling), on
* the reference's own frames kept in tests/golden (img0024/0100/0300.jpg),
  whose decodes are also fixed by SHA-256 in tests/golden/color_golden.json;
* JPEGs PIL encodes here: 4:4:4, 4:2:2, 4:2:0 and grey, odd and tiny sizes
  (chroma widths of 2 or less take libjpeg's box upsampling), quality 1 to
  100, optimised Huffman tables, restart intervals, RGB-coded (Adobe) data.
The product's Huffman decoding (kf_jpeg_coefficients, host code) must give
the oracle's coefficient blocks bit for bit; the device half is checked in
test_gpu_color.py.
"""
import hashlib
import io
import json
import os

import numpy as np
import pytest

from conftest import ROOT

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


def cases(rng, sizes=((480, 640), (23, 37), (17, 5), (9, 3), (1, 1), (2, 7)),
          qualities=(1, 50, 100)):
    for h, w in sizes:
        for mode, subs in (("RGB", (0, 1, 2)), ("L", (None,))):
            for sub in subs:
                for q in qualities:
                    for extra in ({}, {"optimize": True}, {"restart_marker_blocks": 3}):
                        kw = dict(quality=q, **extra)
                        if sub is not None:
                            kw["subsampling"] = sub
                        yield (h, w, mode, sub, q, tuple(extra)), encode(synthetic(rng, h, w, mode), **kw)


def test_reference_frames_match_golden_and_pil(oracle):
    golden = json.load(open(os.path.join(GOLDEN, "color_golden.json")))["frames"]
    for name in FRAMES:
        data = open(os.path.join(GOLDEN, name), "rb").read()
        g = golden[name]
        assert hashlib.sha256(data).hexdigest() == g["file_sha256"]
        got = oracle.jpeg_decode(data)
        assert got is not None and list(got.shape) == g["shape"] == [480, 640, 3]
        assert hashlib.sha256(np.ascontiguousarray(got).tobytes()).hexdigest() == g["bgr_sha256"], name
        for k, v in g["pixels"].items():
            y, x = map(int, k.split(","))
            assert got[y, x].tolist() == v
        np.testing.assert_array_equal(got, pil_bgr(data))


def test_oracle_decode_equals_pil_on_encoded_variants(oracle):
    rng = np.random.default_rng(11)
    n = 0
    for what, data in cases(rng):
        got = oracle.jpeg_decode(data)
        want = pil_bgr(data)
        assert got is not None and got.shape == want.shape, what
        assert (got == want).all(), (what, int((got != want).sum()))
        n += 1
    assert n == 6 * 4 * 3 * 3


def test_oracle_rgb_coded_and_rejections(oracle):
    rng = np.random.default_rng(12)
    img = synthetic(rng, 24, 40)
    data = encode(img, keep_rgb=True, quality=90)
    assert b"Adobe" in data
    np.testing.assert_array_equal(oracle.jpeg_decode(data), pil_bgr(data))
    assert oracle.jpeg_decode(encode(img, progressive=True)) is None
    assert oracle.jpeg_decode(encode(img.convert("CMYK"))) is None
    assert oracle.jpeg_decode(b"\xff\xd8\xff\xd9") is None
    assert oracle.jpeg_decode(b"not a jpeg") is None


def test_host_huffman_decode_equals_oracle_coefficients(kf, oracle):
    """kf_jpeg_coefficients (the product's host half) against the oracle's
    bit-at-a-time decode."""
    for name in FRAMES:
        data = open(os.path.join(GOLDEN, name), "rb").read()
        a, b = kf.jpeg_coefficients(data), oracle.jpeg_coefficients(data)
        assert a.shape == b.shape == (7200, 64)
        np.testing.assert_array_equal(a, b)
        assert kf.jpeg_info(data) == (480, 640, 3)
    rng = np.random.default_rng(13)
    for what, data in cases(rng, sizes=((23, 37), (9, 3), (64, 48)), qualities=(1, 100)):
        a, b = kf.jpeg_coefficients(data), oracle.jpeg_coefficients(data)
        assert a.shape == b.shape, what
        assert (a == b).all(), what


def test_host_decode_rejects_unsupported_streams(kf):
    rng = np.random.default_rng(14)
    img = synthetic(rng, 16, 16)
    for data, msg in ((encode(img, progressive=True), b"progressive"),
                      (encode(img.convert("CMYK")), b"3-component"),
                      (b"\xff\xd8\xff\xd9", b"frame"),
                      (b"GIF89a", b"not a JPEG")):
        with pytest.raises(kf.KinfuError):
            kf.jpeg_info(data) if msg != b"frame" else kf.jpeg_coefficients(data)
        assert msg in kf.lib.kf_last_error()
    # entropy-coded data cut short decodes with zero bits, as libjpeg does
    data = encode(synthetic(rng, 64, 64), quality=90)
    sos = data.index(b"\xff\xda")
    start = sos + 2 + (data[sos + 2] << 8 | data[sos + 3])
    cut = data[:start + (len(data) - start) // 2]
    a, b = kf.jpeg_coefficients(cut), kf.jpeg_coefficients(data)
    assert a.shape == b.shape and (a[:4] == b[:4]).all() and not (a == b).all()


def test_oracle_point_cloud_packing(oracle):
    """threeD_viewer.cpp:20-55: rgb = r << 16 | g << 8 | b as float bits where
    z != 0 (-0.0 counts as zero), zeros elsewhere."""
    rng = np.random.default_rng(15)
    H, W = 6, 9
    x, y, z = (rng.normal(size=(H, W)).astype(np.float32) * 100 for _ in range(3))
    z[::2, 1::3] = 0.0
    z[1, 1] = -0.0
    bgr = rng.integers(0, 256, (H, W, 3)).astype(np.uint8)
    got = oracle.point_cloud(x, y, z, bgr)
    bits = got[..., 3].view(np.uint32)
    keep = z != 0
    want = (bgr[..., 2].astype(np.uint32) << 16) | (bgr[..., 1].astype(np.uint32) << 8) | bgr[..., 0]
    np.testing.assert_array_equal(bits[keep], want[keep])
    np.testing.assert_array_equal(got[keep][:, :3], np.stack([x, y, z], -1)[keep])
    assert (got[~keep].view(np.uint32) == 0).all()
