"""Write the golden fixtures in tests/golden/ (run in the build container).

Inputs come from the SplitMix64 `signed_unit` stream of the reference's tests
(ctest/flagfft_test.h:390-417) with the seeds recorded in each fixture; outputs
are numpy 2.2 pocketfft results in complex128 / float64, each cross-checked
against the long-double O(N^2) DFT of oracle.c before it is written.  The
fixtures therefore pin the oracle to an independent FFT implementation.

Usage: python tests/golden/make_golden.py
"""
from __future__ import annotations

import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))

from oracle import oracle as ora  # noqa: E402

C2C_SIZES = [16, 23, 64, 67, 81, 243, 361, 997, 1009, 4096, 10007, 10935]
INV_SIZES = [23, 256, 997]
R2C_SIZES = [15, 16, 67, 6000]
C2R_SIZES = [15, 16, 6000]
TWO_D = [(16, 16), (23, 23), (64, 32)]
TYPE_C2C, TYPE_R2C, TYPE_C2R = 0x29, 0x2A, 0x2C


def _check(out, ref, what):
    err = np.linalg.norm(out - ref) / max(np.linalg.norm(ref), 1e-300)
    assert err < 1e-13, (what, err)


def main():
    fx = {}
    for n in C2C_SIZES:
        seed = ora.accuracy_seed(TYPE_C2C, n, 1, 7)
        x = ora.complex_signed_unit(seed, (n,))
        out = np.fft.fft(x)
        _check(out, ora.dft(x), f"c2c {n}")
        fx[f"c2c_fwd_{n}_seed"] = np.uint64(seed)
        fx[f"c2c_fwd_{n}"] = out
    for n in INV_SIZES:
        seed = ora.accuracy_seed(TYPE_C2C, n, 1, 8)
        x = ora.complex_signed_unit(seed, (n,))
        out = np.fft.ifft(x) * n
        _check(out, ora.dft(x, sign=1), f"c2c inv {n}")
        fx[f"c2c_inv_{n}_seed"] = np.uint64(seed)
        fx[f"c2c_inv_{n}"] = out
    for n in R2C_SIZES:
        seed = ora.accuracy_seed(TYPE_R2C, n, 1, 7)
        x = ora.signed_unit(seed, n)
        out = np.fft.rfft(x)
        _check(out, ora.dft(x.astype(np.complex128))[: n // 2 + 1], f"r2c {n}")
        fx[f"r2c_{n}_seed"] = np.uint64(seed)
        fx[f"r2c_{n}"] = out
    for n in C2R_SIZES:
        seed = ora.accuracy_seed(TYPE_C2R, n, 1, 7)
        X = ora.complex_signed_unit(seed, (n // 2 + 1,))  # DC / Nyquist imaginary parts left non-zero
        out = np.fft.irfft(X, n) * n                       # pocketfft drops them, like the reference
        Xh = X.copy()
        Xh[0] = Xh[0].real
        if n % 2 == 0:
            Xh[-1] = Xh[-1].real
        full = np.concatenate([Xh, np.conj(Xh[1:(n + 1) // 2][::-1])])
        _check(out, ora.dft(full, sign=1).real, f"c2r {n}")
        fx[f"c2r_{n}_seed"] = np.uint64(seed)
        fx[f"c2r_{n}"] = out
    for n0, n1 in TWO_D:
        seed = ora.accuracy_seed(TYPE_C2C, n0 * 1000 + n1, 1, 9)
        x = ora.complex_signed_unit(seed, (n0, n1))
        out = np.fft.fft2(x)
        rows = ora.dft(x)
        _check(out, ora.dft(np.ascontiguousarray(rows.T)).T, f"2d {n0}x{n1}")
        fx[f"fft2_{n0}x{n1}_seed"] = np.uint64(seed)
        fx[f"fft2_{n0}x{n1}"] = out
        seed_r = ora.accuracy_seed(TYPE_R2C, n0 * 1000 + n1, 1, 9)
        xr = ora.signed_unit(seed_r, n0 * n1).reshape(n0, n1)
        fx[f"rfft2_{n0}x{n1}_seed"] = np.uint64(seed_r)
        fx[f"rfft2_{n0}x{n1}"] = np.fft.rfft2(xr)
    np.savez_compressed(os.path.join(HERE, "golden_fft.npz"), **fx)
    print(f"wrote {len(fx)} arrays to golden_fft.npz")


if __name__ == "__main__":
    main()
