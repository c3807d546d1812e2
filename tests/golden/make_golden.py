"""Writes the golden digests under tests/golden/ from the reference's data
files (run in a container that has the reference checkout; the digests are
committed, the reference files are not).

  python tests/golden/make_golden.py /path/to/rusterizer
"""
import hashlib
import os
import sys

import numpy as np
from PIL import Image

HERE = os.path.dirname(os.path.abspath(__file__))


def main(ref: str):
    png = os.path.join(ref, "images", "checkerboard.png")
    rgba = np.array(Image.open(png).convert("RGBA"))
    digest = hashlib.sha256(np.ascontiguousarray(rgba).tobytes()).hexdigest()
    with open(os.path.join(HERE, "checkerboard_png_rgba_sha256.txt"), "w") as f:
        f.write(f"{digest}  images/checkerboard.png decoded to {rgba.shape[1]}x{rgba.shape[0]} RGBA8\n")
    print(digest)


if __name__ == "__main__":
    main(sys.argv[1] if len(sys.argv) > 1 else "/root/reference")
