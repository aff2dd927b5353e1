"""Writes tests/golden/color_golden.json: for each reference colour frame
kept here (img0024.jpg, img0100.jpg, img0300.jpg, copied unchanged from the
reference's data/ directory), the SHA-256 of the H x W x 3 B, G, R image
that libjpeg-turbo's default decompression gives (through PIL, with the
channels reversed to OpenCV's order), its size, and a few pixels.

cv::imread(path) (system_utility.cpp:70-75) decodes with the same library
and defaults, so these are the images the reference's LoadColorData holds.
Regenerate only on purpose:

    python tests/golden/make_color_golden.py
"""
import hashlib
import io
import json
import os

import numpy as np
from PIL import Image, features

HERE = os.path.dirname(os.path.abspath(__file__))
FRAMES = ("img0024.jpg", "img0100.jpg", "img0300.jpg")


def pil_bgr(data):
    return np.ascontiguousarray(np.asarray(Image.open(io.BytesIO(data)).convert("RGB"))[:, :, ::-1])


def main():
    out = {"decoder": f"PIL {Image.__version__}, libjpeg-turbo "
                      f"{features.version('libjpeg_turbo')}", "frames": {}}
    for name in FRAMES:
        data = open(os.path.join(HERE, name), "rb").read()
        bgr = pil_bgr(data)
        out["frames"][name] = {
            "file_sha256": hashlib.sha256(data).hexdigest(),
            "shape": list(bgr.shape),
            "bgr_sha256": hashlib.sha256(bgr.tobytes()).hexdigest(),
            "pixels": {f"{y},{x}": bgr[y, x].tolist()
                       for y, x in ((0, 0), (0, 639), (240, 320), (479, 0), (479, 639), (123, 457))},
        }
    with open(os.path.join(HERE, "color_golden.json"), "w") as f:
        json.dump(out, f, indent=1)
        f.write("\n")


if __name__ == "__main__":
    main()
