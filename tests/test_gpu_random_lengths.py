"""Seeded sweep over lengths the other tests do not name (GPU).

Sixty lengths drawn log-uniformly from [2, 2^21] (a fixed seed, so the set is
the same on every run) with batches from {1, 2, 3, 7}: every plan kind the
planner picks for them (leaves, split leaves, register-resident and LDS
chains with partial tiles, Rader, Bluestein in LDS and as chains) against
the FP64 oracle within the stated bounds, C2C both directions and Z2Z
forward on every sequence, and for even lengths the R2C -> C2R round trip.
"""
import numpy as np
import pytest

from flagfft_amd import flagfft as ff
from oracle import oracle as ora

from .helpers import check_accuracy, complex_input, dev_tensor, host, real_input, run_c2c, stream_of, sync

pytestmark = pytest.mark.gpu


def _cases():
    rng = np.random.default_rng(20261203)
    out = []
    for _ in range(60):
        n = int(round(float(np.exp(rng.uniform(np.log(2), np.log(1 << 21))))))
        b = int(rng.choice([1, 2, 3, 7]))
        out.append((max(n, 2), b))
    return out


CASES = _cases()


@pytest.mark.parametrize("n,batch", CASES, ids=[f"{n}x{b}" for n, b in CASES])
def test_complex_lengths(gpu, n, batch):
    x = complex_input(n * 31 + batch, (batch, n), np.complex64, "normal")
    for direction in (ff.FORWARD, ff.INVERSE):
        y = run_c2c(gpu, x, direction, n, batch)
        check_accuracy(y, ora.fft(x.astype(np.complex128), sign=direction), ff.Type.C2C, n, row_len=n,
                       what=f"C2C n={n} batch={batch} dir={direction}")
    if n <= (1 << 18):
        xz = complex_input(n * 37 + batch, (batch, n), np.complex128, "normal")
        y = run_c2c(gpu, xz, ff.FORWARD, n, batch)
        check_accuracy(y, ora.fft(xz), ff.Type.Z2Z, n, row_len=n, what=f"Z2Z n={n} batch={batch}")


@pytest.mark.parametrize("n,batch", [(n + (n & 1), b) for n, b in CASES[::3]],
                         ids=[f"{n + (n & 1)}x{b}" for n, b in CASES[::3]])
def test_real_roundtrip_lengths(gpu, n, batch):
    import torch

    x = real_input(n * 41 + batch, (batch, n), np.float32)
    xd = dev_tensor(x, gpu)
    h = n // 2 + 1
    yd = torch.zeros((batch, h), dtype=torch.complex64, device=gpu)
    zd = torch.zeros_like(xd)
    with ff.Plan.one_d(n, ff.Type.R2C, batch) as pf, ff.Plan.one_d(n, ff.Type.C2R, batch) as pi:
        for p in (pf, pi):
            p.set_stream(stream_of(gpu))
        pf.r2c(xd.data_ptr(), yd.data_ptr())
        pi.c2r(yd.data_ptr(), zd.data_ptr())
        sync(gpu)
    check_accuracy(host(yd), ora.rfft(x.astype(np.float64)), ff.Type.R2C, n, row_len=h, what=f"R2C n={n}")
    check_accuracy(host(zd), n * x.astype(np.float64), ff.Type.C2R, n, row_len=n, factor=2.0,
                   what=f"C2R roundtrip n={n}")
