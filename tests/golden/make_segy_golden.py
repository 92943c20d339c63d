"""Write the SEG-Y golden fixture with the reference's own WriteSGY
(SGYWrite.cpp:3-55, compiled in oracle/_ref).

Inputs: a seeded 12-trace x 40-sample image (values spanning signs, zero,
tiny and large magnitudes) and the reference's geometry arrays
(kernel.cu:1197-1204).  Outputs in this directory:
  segy_input.npy     the image, float32 [trace][sample]
  segy_expected.sgy  the bytes WriteSGY produces from SGY_Model.sgy
Run from the repository root: python tests/golden/make_segy_golden.py
"""
import os
import shutil
import sys
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(HERE))
import oracle_bind as ob  # noqa: E402

NTR, NS, DZ = 12, 40, 10


def golden_input():
    rng = np.random.default_rng(20261202)
    img = rng.normal(0, 1, (NTR, NS)).astype(np.float32)
    img *= np.logspace(-30, 30, NS, dtype=np.float64).astype(np.float32)[None, :]
    img[0, :5] = 0.0
    img[1, 0] = -118.625
    img[1, 1] = 1.0
    img[2, 2] = np.float32(3.4e38)
    img[2, 3] = np.float32(-1e-38)
    return img


def geometry(n):
    sx = np.array([1.0 * i * 10 for i in range(n)], np.float32)
    sy = np.array([-1.0 * i * 10 for i in range(n)], np.float32)
    dsr = np.array([1000 - i for i in range(n)], np.float32)
    return sx, sy, dsr


def main():
    if not ob.ref_available():
        raise SystemExit("oracle/_ref/libref.so not built")
    img = golden_input()
    sx, sy, dsr = geometry(NTR)
    with tempfile.TemporaryDirectory() as d:
        shutil.copy(os.path.join(HERE, "SGY_Model.sgy"), os.path.join(d, "SGY_Model.sgy"))
        cwd = os.getcwd()
        os.chdir(d)
        try:
            ob.ref().ref_write_sgy(np.ascontiguousarray(img).ravel(), NTR, NS, DZ, sx, sy, 1.0, 1.0,
                                   dsr, b"out.sgy")
        finally:
            os.chdir(cwd)
        shutil.copy(os.path.join(d, "out.sgy"), os.path.join(HERE, "segy_expected.sgy"))
    np.save(os.path.join(HERE, "segy_input.npy"), img)
    print("wrote", os.path.join(HERE, "segy_expected.sgy"))


if __name__ == "__main__":
    main()
