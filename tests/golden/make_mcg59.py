#!/usr/bin/env python3
"""Writes the random-stream golden vectors: draws of Intel MKL's VSL MCG59 generator, the one the
reference's subroutines.f90:16, 49-159 uses (vslNewStreamEx with two seed words, vdRngUniform
with VSL_RNG_METHOD_UNIFORM_STD, vdRngGaussian with VSL_RNG_METHOD_GAUSSIAN_BOXMULLER,
vslSkipAheadStream), taken from the MKL runtime library installed with this image's conda
(libmkl_rt, called through ctypes; the version string goes into the file).  They pin the
oracle's restatement of the generator (x_{n+1} = 13^13 x_n mod 2^59, u = x / 2^59, Box-Muller on
consecutive pairs) and of the skip-ahead the reference uses to place its draws.

    python tests/golden/make_mcg59.py      # rewrites mcg59_mkl.npz beside this file
"""
import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
MKL = "/opt/conda/lib/libmkl_rt.so"
VSL_BRNG_MCG59 = 0x400000
SEEDS = ((1, 2), (12345, 678), (4294967295, 4294967295), (7, 0), (0, 0))
SKIPS = (0, 1, 1000000007, 4 * 128 ** 3 + 12, 4004 * 98765 + 8)
NDRAW = 64


def main():
    mkl = C.CDLL(MKL)
    buf = C.create_string_buffer(256)
    mkl.mkl_get_version_string(buf, 256)
    uni = np.zeros((len(SEEDS), len(SKIPS), NDRAW))
    gau = np.zeros((len(SEEDS), len(SKIPS), NDRAW))
    for i, (w0, w1) in enumerate(SEEDS):
        for j, skip in enumerate(SKIPS):
            for out, fn in ((uni, "vdRngUniform"), (gau, "vdRngGaussian")):
                st = C.c_void_p()
                seed = (C.c_uint * 2)(w0, w1)
                assert mkl.vslNewStreamEx(C.byref(st), C.c_int(VSL_BRNG_MCG59), C.c_int(2), seed) == 0
                assert mkl.vslSkipAheadStream(st, C.c_longlong(skip)) == 0
                row = np.zeros(NDRAW)
                rc = getattr(mkl, fn)(C.c_int(0), st, C.c_int(NDRAW), row.ctypes.data_as(C.c_void_p),
                                      C.c_double(0.0), C.c_double(1.0))
                assert rc == 0
                out[i, j] = row
                mkl.vslDeleteStream(C.byref(st))
    np.savez(os.path.join(HERE, "mcg59_mkl.npz"), seeds=np.array(SEEDS, np.uint32), skips=np.array(SKIPS, np.uint64),
             uniform=uni, gauss=gau, mkl_version=np.array(buf.value.decode().strip()))


if __name__ == "__main__":
    main()
