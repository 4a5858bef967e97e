"""Writes the golden fixtures used by tests/test_oracle.py and the GPU parity tests.

* imfill_bw1.npz: the 8x8 BW1 matrix and BW2 = imfill(BW1,[3 3],8) from the MATLAB imfill
  documentation, quoted in the reference's src/segment/test/src/imfillTest.cpp:26-39 (the expected
  BW2 is the documented MATLAB output).
* size_phantom_1024.npz: the disk phantom of the reference's src/segment/test/sizePhantom.m
  (disks of radius r, (r+1)/2-1, ... placed along the diagonal, one pixel apart), generated at
  1024^2 with rmax = 256 instead of 4096^2 / 1024 to keep the fixture small.
* watershed.png: copied unchanged from the reference's src/segment/test/watershed.png (an input
  image, 256x256 RGB; the reference holds no expected output for it).
"""
import os
import shutil

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))

BW1 = np.array([[1, 0, 0, 0, 0, 0, 0, 0],
                [1, 1, 1, 1, 1, 0, 0, 0],
                [1, 0, 0, 0, 1, 0, 1, 0],
                [1, 0, 0, 0, 1, 1, 1, 0],
                [1, 1, 1, 1, 0, 1, 1, 1],
                [1, 0, 0, 1, 1, 0, 1, 0],
                [1, 0, 0, 0, 1, 0, 1, 0],
                [1, 0, 0, 0, 1, 1, 1, 0]], np.uint8)
BW2 = np.array([[1, 0, 0, 0, 0, 0, 0, 0],
                [1, 1, 1, 1, 1, 0, 0, 0],
                [1, 1, 1, 1, 1, 0, 1, 0],
                [1, 1, 1, 1, 1, 1, 1, 0],
                [1, 1, 1, 1, 1, 1, 1, 1],
                [1, 0, 0, 1, 1, 1, 1, 0],
                [1, 0, 0, 0, 1, 1, 1, 0],
                [1, 0, 0, 0, 1, 1, 1, 0]], np.uint8)


def size_phantom(n=1024, rmax=256):
    z = np.zeros((n, n), np.uint8)
    r = rmax - 1
    start = 0
    radii = []
    while r >= 0:
        d = 2 * r + 1
        i = np.arange(-r, r + 1)
        ii, jj = np.meshgrid(i, i, indexing="ij")
        z[start:start + d, start:start + d] = ((ii * ii + jj * jj) <= r * r) * 255
        radii.append(r)
        start = start + d + 1
        r = (r + 1) // 2 - 1
    return z, np.array(radii)


if __name__ == "__main__":
    np.savez_compressed(os.path.join(HERE, "imfill_bw1.npz"), bw1=BW1 * 255, bw2=BW2 * 255)
    img, radii = size_phantom()
    np.savez_compressed(os.path.join(HERE, "size_phantom_1024.npz"), img=img, radii=radii)
    src = "/root/reference/src/segment/test/watershed.png"
    if os.path.exists(src):
        shutil.copyfile(src, os.path.join(HERE, "watershed.png"))
    print("fixtures written to", HERE)
