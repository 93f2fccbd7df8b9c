"""Writes the small golden fixtures used by the CPU tests.

fft_golden.npz: numpy.fft.rfftn / irfftn of seeded inputs (FFTW conventions:
forward e^{-2 pi i}, unnormalised) in the reference's Fortran layout
(arrays are (nz, ny, nx) C-order == f(nx, ny, nz)), including odd-radix and
anisotropic shapes.  The reference repository holds no golden vectors of its
own for this path, so these library outputs are the FFT pin.

Run:  python tests/golden/make_golden.py
"""
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
SHAPES = [(8, 8, 8), (6, 5, 4), (16, 12, 10), (24, 24, 24), (48, 8, 8), (8, 8, 48), (12, 16, 20)]


def main():
    rng = np.random.default_rng(20240601)
    out = {}
    for nz, ny, nx in SHAPES:
        f = rng.standard_normal((nz, ny, nx))
        key = f"{nx}x{ny}x{nz}"
        out[f"in_{key}"] = f
        out[f"r2c_{key}"] = np.fft.rfftn(f)
        G = rng.standard_normal((nz, ny, nx // 2 + 1)) + 1j * rng.standard_normal((nz, ny, nx // 2 + 1))
        out[f"cin_{key}"] = G
        out[f"c2r_{key}"] = np.fft.irfftn(G, s=(nz, ny, nx), axes=(0, 1, 2)) * (nx * ny * nz)
    np.savez_compressed(os.path.join(HERE, "fft_golden.npz"), **out)


if __name__ == "__main__":
    main()
