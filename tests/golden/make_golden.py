#!/usr/bin/env python3
"""Writes the FFT golden vectors: seeded complex inputs and numpy's unnormalized transform
with the sign of FFTW_BACKWARD (sum_x f(x) e^{+2 pi i k.x / N}, grav_waves.f90:117-118),
i.e. N^3 * numpy.fft.ifftn.  numpy's pocketfft is independent of the oracle's mixed-radix
code, so these pin the oracle's transform (and through it the GPU's).

    python tests/golden/make_golden.py      # rewrites fft_<N>.npz beside this file
"""
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))


def main():
    for N, seed in ((12, 1201), (16, 1601)):
        rng = np.random.default_rng(seed)
        f = rng.standard_normal((N, N, N)) + 1j * rng.standard_normal((N, N, N))
        out = np.fft.ifftn(f) * N ** 3
        np.savez(os.path.join(HERE, f"fft_{N}.npz"), input=f, backward=out)


if __name__ == "__main__":
    main()
