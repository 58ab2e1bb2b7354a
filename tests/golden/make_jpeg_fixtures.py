"""Writes the JPEG fixtures of tests/golden/jpeg/: small frames encoded by
Pillow (libjpeg-turbo) at several sizes, subsamplings and restart intervals,
with the RGB that Pillow's decoder (libjpeg defaults: ISLOW IDCT, fancy
upsampling) returns for them.  The reference's reader (readjpeg/src/main/cpp/
native.cpp:59-134) is libjpeg with its defaults too, so these are its
expected outputs.  Run: python tests/golden/make_jpeg_fixtures.py"""
import io
import os

import numpy as np
from PIL import Image

OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)), "jpeg")
CASES = [  # name, (w, h), save options
    ("w14h8_444_q75", (14, 8), dict(quality=75, subsampling=0)),
    ("w5h3_422_q90", (5, 3), dict(quality=90, subsampling=1)),
    ("w37h29_420_q95", (37, 29), dict(quality=95, subsampling=2)),
    ("w4h3_420_q85", (4, 3), dict(quality=85, subsampling=2)),
    ("w40h24_420_rst3_opt", (40, 24), dict(quality=80, subsampling=2, restart_marker_blocks=3, optimize=True)),
]


def frame(w, h, seed):
    rng = np.random.default_rng(seed)
    y, x = np.mgrid[0:h, 0:w]
    img = np.stack([(x * 7 + y * 3) % 256, (x * x + y) % 256, 128 + 60 * np.sin(x / 5.0) * np.cos(y / 7.0)], -1)
    return (img + rng.normal(0, 12, img.shape)).clip(0, 255).astype(np.uint8)


def main():
    os.makedirs(OUT, exist_ok=True)
    for k, (name, (w, h), opts) in enumerate(CASES):
        buf = io.BytesIO()
        Image.fromarray(frame(w, h, 100 + k)).save(buf, "JPEG", **opts)
        data = buf.getvalue()
        with open(os.path.join(OUT, name + ".jpg"), "wb") as f:
            f.write(data)
        rgb = np.asarray(Image.open(io.BytesIO(data)).convert("RGB"))
        np.save(os.path.join(OUT, name + ".rgb.npy"), rgb)
        print(name, len(data), "bytes", rgb.shape)


if __name__ == "__main__":
    main()
