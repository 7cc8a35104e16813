"""R2C / C2R / D2Z / Z2D parity against the oracle through the C ABI (GPU).

Mirrors ctest/test_exec_{r2c,c2r,d2z,z2d}_*.cpp and the roundtrip tests
ctest/test_exec_{r2c_c2r,d2z_z2d}_*.cpp: half-spectrum outputs, C2R ignoring
Im of DC/Nyquist, zero input -> exact zero, R2C -> C2R = N x, padded in-place layouts.
"""
import numpy as np
import pytest

from flagfft_amd import flagfft as ff
from oracle import oracle as ora

from .helpers import check_accuracy, complex_input, dev_tensor, host, real_input, stream_of, sync

pytestmark = pytest.mark.gpu

# even lengths whose half is a Rader (194, 2018) or Bluestein (9998, 65538) prime-factor core
SIZES = [2, 3, 15, 16, 64, 67, 81, 194, 243, 256, 361, 997, 1009, 2018, 4096, 6000, 8192, 9998, 10007, 16384, 32768,
         65536 + 2,
         # halves on the 3 / 5 / 15 x 2^k and 6000 / 10000-point leaves
         3072, 2560, 3840, 4800, 12000, 20000]


def _r2c(dev, x, n, batch, dbl):
    t = ff.Type.D2Z if dbl else ff.Type.R2C
    xd = dev_tensor(x, dev)
    yd = dev_tensor(np.zeros((batch, n // 2 + 1), np.complex128 if dbl else np.complex64), dev)
    with ff.Plan.one_d(n, t, batch) as p:
        p.set_stream(stream_of(dev))
        (p.d2z if dbl else p.r2c)(xd.data_ptr(), yd.data_ptr())
        sync(dev)
    return host(yd)


def _c2r(dev, X, n, batch, dbl):
    t = ff.Type.Z2D if dbl else ff.Type.C2R
    xd = dev_tensor(X, dev)
    keep = xd.clone()
    yd = dev_tensor(np.zeros((batch, n), np.float64 if dbl else np.float32), dev)
    with ff.Plan.one_d(n, t, batch) as p:
        p.set_stream(stream_of(dev))
        (p.z2d if dbl else p.c2r)(xd.data_ptr(), yd.data_ptr())
        sync(dev)
    assert bool((xd == keep).all()), "C2R must not modify its input"
    return host(yd)


@pytest.mark.parametrize("dbl", [False, True], ids=["fp32", "fp64"])
@pytest.mark.parametrize("n", SIZES)
def test_r2c(gpu, n, dbl):
    t = ff.Type.D2Z if dbl else ff.Type.R2C
    for batch in (1, 3):
        x = real_input(ora.accuracy_seed(int(t), n, batch), (batch, n), np.float64 if dbl else np.float32)
        y = _r2c(gpu, x, n, batch, dbl)
        check_accuracy(y, ora.rfft(x.astype(np.float64)), t, n, row_len=n // 2 + 1, what=f"batch={batch}")


@pytest.mark.parametrize("dbl", [False, True], ids=["fp32", "fp64"])
@pytest.mark.parametrize("n", SIZES)
def test_c2r(gpu, n, dbl):
    t = ff.Type.Z2D if dbl else ff.Type.C2R
    for batch in (1, 3):
        X = complex_input(ora.accuracy_seed(int(t), n, batch), (batch, n // 2 + 1),
                          np.complex128 if dbl else np.complex64)
        # DC and Nyquist keep non-zero imaginary parts: both sides must ignore them
        y = _c2r(gpu, X, n, batch, dbl)
        check_accuracy(y, ora.irfft(X.astype(np.complex128), n), t, n, row_len=n, what=f"batch={batch}")


@pytest.mark.parametrize("n", [256, 6000, 997, 32768])
def test_roundtrip_r2c_c2r(gpu, n):
    """R2C -> C2R = N x (ctest/test_exec_r2c_c2r_ct_b.cpp:54-70)."""
    for dbl, tf, ti in ((False, ff.Type.R2C, ff.Type.C2R), (True, ff.Type.D2Z, ff.Type.Z2D)):
        x = real_input(ora.accuracy_seed(int(tf), n, 4), (4, n), np.float64 if dbl else np.float32)
        X = _r2c(gpu, x, n, 4, dbl)
        z = _c2r(gpu, X, n, 4, dbl)
        check_accuracy(z, x.astype(np.float64) * n, ti, n, row_len=n, factor=2.0, what="roundtrip")


def test_config_6000x128_roundtrips(gpu):
    """R2C->C2R FP32 and D2Z->Z2D FP64 at N=6000 batch 128, inputs U[-1,1]."""
    n, batch = 6000, 128
    for dbl, tf, ti in ((False, ff.Type.R2C, ff.Type.C2R), (True, ff.Type.D2Z, ff.Type.Z2D)):
        x = real_input(0x6000, (batch, n), np.float64 if dbl else np.float32)
        X = _r2c(gpu, x, n, batch, dbl)
        check_accuracy(X, ora.rfft(x.astype(np.float64)), tf, n, row_len=n // 2 + 1, what="config fwd")
        z = _c2r(gpu, X, n, batch, dbl)
        check_accuracy(z, x.astype(np.float64) * n, ti, n, row_len=n, factor=2.0, what="config roundtrip")


@pytest.mark.parametrize("t", [ff.Type.C2R, ff.Type.Z2D])
def test_zero_input_is_exact(gpu, t):
    """ctest/test_exec_c2r_ct_s.cpp:76-96 and the Z2D twin."""
    dbl = t == ff.Type.Z2D
    for n in (256, 255, 32768, 67):
        X = np.zeros((1, n // 2 + 1), np.complex128 if dbl else np.complex64)
        y = _c2r(gpu, X, n, 1, dbl)
        assert np.all(y == 0.0) and np.all(np.isfinite(y))


@pytest.mark.parametrize("n", [256, 6000, 243, 67])
def test_padded_in_place_r2c_c2r(gpu, n):
    """In-place real transforms with padded rows of 2(N/2+1) reals (src/exec/c_api_internal.cpp:74-89)."""
    import torch

    h = n // 2 + 1
    batch = 3
    x = real_input(n + 5, (batch, n), np.float32)
    buf = torch.zeros((batch, 2 * h), dtype=torch.float32, device=gpu)
    buf[:, :n] = torch.from_numpy(x).to(gpu)
    with ff.Plan.many(1, [n], None, 1, 2 * h, None, 1, h, ff.Type.R2C, batch) as p:
        p.set_stream(stream_of(gpu))
        p.r2c(buf.data_ptr(), buf.data_ptr())
        sync(gpu)
    X = host(buf).view(np.complex64)
    check_accuracy(X, ora.rfft(x.astype(np.float64)), ff.Type.R2C, n, row_len=h, what="in-place R2C")
    with ff.Plan.many(1, [n], None, 1, h, None, 1, 2 * h, ff.Type.C2R, batch) as p:
        p.set_stream(stream_of(gpu))
        p.c2r(buf.data_ptr(), buf.data_ptr())
        sync(gpu)
    z = host(buf)[:, :n]
    check_accuracy(z, x.astype(np.float64) * n, ff.Type.C2R, n, row_len=n, factor=2.0, what="in-place C2R")


def test_c2r_large_batch_multiple_seeds(gpu):
    """ctest/test_exec_c2r_ct_b.cpp:76-117: N=16384, batch 256, seeds 1..10, Im of DC / Nyquist zeroed."""
    n, batch = 16384, 256
    h = n // 2 + 1
    for seed in range(1, 11):
        X = complex_input(ora.accuracy_seed(int(ff.Type.C2R), n, batch, seed), (batch, h), np.complex64)
        X[:, 0] = X[:, 0].real
        X[:, n // 2] = X[:, n // 2].real
        y = _c2r(gpu, X, n, batch, False)
        check_accuracy(y, ora.irfft(X.astype(np.complex128), n), ff.Type.C2R, n, row_len=n, what=f"seed={seed}")


# batch 256 of the reference's "_b" binaries (ctest/flagfft_test.h:509-526)
BATCH_SIZES = [16, 64, 81, 243, 256, 361, 512, 2048, 4096]


@pytest.mark.parametrize("dbl", [False, True], ids=["r2c-c2r", "d2z-z2d"])
@pytest.mark.parametrize("n", BATCH_SIZES)
def test_real_batch_256(gpu, n, dbl):
    tf, ti = (ff.Type.D2Z, ff.Type.Z2D) if dbl else (ff.Type.R2C, ff.Type.C2R)
    batch = 256
    x = real_input(ora.accuracy_seed(int(tf), n, batch), (batch, n), np.float64 if dbl else np.float32)
    X = _r2c(gpu, x, n, batch, dbl)
    check_accuracy(X, ora.rfft(x.astype(np.float64)), tf, n, row_len=n // 2 + 1, what="R2C batch 256")
    Y = complex_input(ora.accuracy_seed(int(ti), n, batch), (batch, n // 2 + 1), np.complex128 if dbl else np.complex64)
    y = _c2r(gpu, Y, n, batch, dbl)
    check_accuracy(y, ora.irfft(Y.astype(np.complex128), n), ti, n, row_len=n, what="C2R batch 256")


@pytest.mark.parametrize("scale", [2.0 ** -20, 2.0 ** 20])
@pytest.mark.parametrize("n", [64, 997, 4096, 6000])
def test_real_input_scales(gpu, n, scale):
    """Inputs scaled by 2^-20 and 2^20 (ctest/flagfft_test.h:457), both precisions."""
    for dbl, tf, ti in ((False, ff.Type.R2C, ff.Type.C2R), (True, ff.Type.D2Z, ff.Type.Z2D)):
        x = real_input(ora.accuracy_seed(int(tf), n, 4), (4, n), np.float64 if dbl else np.float32, scale=scale)
        X = _r2c(gpu, x, n, 4, dbl)
        check_accuracy(X, ora.rfft(x.astype(np.float64)), tf, n, row_len=n // 2 + 1, what=f"scale={scale}")
        Y = complex_input(ora.accuracy_seed(int(ti), n, 4), (4, n // 2 + 1), np.complex128 if dbl else np.complex64,
                          scale=scale)
        y = _c2r(gpu, Y, n, 4, dbl)
        check_accuracy(y, ora.irfft(Y.astype(np.complex128), n), ti, n, row_len=n, what=f"C2R scale={scale}")
