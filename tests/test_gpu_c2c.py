"""C2C / Z2Z parity against the oracle through the C ABI (GPU).

Mirrors the reference's 1D accuracy suite (ctest/test_exec_{c2c,z2z}_{fwd,inv}_{ct,bs}_{s,b}.cpp):
its sizes, batches, input scales and roundtrip checks, with the FP64 oracle in place of cuFFT.
"""
import numpy as np
import pytest

from flagfft_amd import flagfft as ff
from oracle import oracle as ora

from .helpers import check_accuracy, complex_input, run_c2c

pytestmark = pytest.mark.gpu

CT_SIZES = [1, 2, 3, 5, 7, 11, 13, 16, 64, 81, 243, 256, 361, 512, 1024, 2048, 4096, 8192, 16384]
BS_SIZES = [17, 23, 67, 97, 997, 1009, 4099]
EXTRA = [6000, 10935, 10007, 3 * 5 * 7 * 11 * 13, 2 * 4999, 20000, 65536, 1 << 20]


def _case(dev, n, batch, dtype, direction, scale=1.0, kind="uniform", inplace=False):
    t = ff.Type.Z2Z if dtype == np.complex128 else ff.Type.C2C
    seed = ora.accuracy_seed(int(t), n, batch, 0)
    x = complex_input(seed, (batch, n), dtype, kind, scale)
    y = run_c2c(dev, x, direction, n, batch, inplace=inplace)
    ref = ora.fft(x.astype(np.complex128), sign=direction)
    return check_accuracy(y, ref, t, n, row_len=n, what=f"dir={direction} batch={batch} scale={scale}")


@pytest.mark.parametrize("dtype", [np.complex64, np.complex128], ids=["c2c", "z2z"])
@pytest.mark.parametrize("n", CT_SIZES)
def test_cooley_tukey_sizes(gpu, n, dtype):
    for batch in (1, 4):
        for direction in (ff.FORWARD, ff.INVERSE):
            _case(gpu, n, batch, dtype, direction)


@pytest.mark.parametrize("dtype", [np.complex64, np.complex128], ids=["c2c", "z2z"])
@pytest.mark.parametrize("n", BS_SIZES)
def test_prime_sizes(gpu, n, dtype):
    for batch in (1, 4):
        for direction in (ff.FORWARD, ff.INVERSE):
            _case(gpu, n, batch, dtype, direction)


@pytest.mark.parametrize("dtype", [np.complex64, np.complex128], ids=["c2c", "z2z"])
@pytest.mark.parametrize("n", EXTRA)
def test_multipass_and_bluestein_sizes(gpu, n, dtype):
    batch = 2 if n < (1 << 20) else 1
    for direction in (ff.FORWARD, ff.INVERSE):
        _case(gpu, n, batch, dtype, direction)


@pytest.mark.parametrize("dtype", [np.complex64, np.complex128], ids=["c2c", "z2z"])
@pytest.mark.parametrize("n", [16, 64, 81, 243, 256, 361, 512, 2048, 4096, 8192, 16384, 23, 67, 997, 1009])
def test_batch_256(gpu, n, dtype):
    """Batch 256 of the reference's "_b" binaries (ctest/flagfft_test.h:509-526), both directions."""
    for direction in (ff.FORWARD, ff.INVERSE):
        _case(gpu, n, 256, dtype, direction)


@pytest.mark.parametrize("scale", [2.0 ** -20, 2.0 ** 20])
@pytest.mark.parametrize("n", [64, 997, 4096, 16384])
def test_input_scales(gpu, n, scale):
    """ctest/flagfft_test.h:457 scales inputs by 2^-20, 1, 2^20."""
    _case(gpu, n, 4, np.complex64, ff.FORWARD, scale=scale)


@pytest.mark.parametrize("n", [16, 4096, 16384, 67, 997, 10007, 10935])
def test_roundtrip_is_n_times_x(gpu, n):
    """inverse(forward(x)) = N x (ctest/test_exec_c2c_inv_bs_s.cpp:73-91)."""
    for dtype, t in ((np.complex64, ff.Type.C2C), (np.complex128, ff.Type.Z2Z)):
        x = complex_input(ora.accuracy_seed(int(t), n, 4, 1), (4, n), dtype)
        y = run_c2c(gpu, x, ff.FORWARD, n, 4)
        z = run_c2c(gpu, y, ff.INVERSE, n, 4)
        check_accuracy(z, x.astype(np.complex128) * n, t, n, row_len=n, factor=2.0, what="roundtrip")


@pytest.mark.parametrize("n", [256, 4096, 16384, 67, 997, 10007])
def test_in_place(gpu, n):
    for dtype in (np.complex64, np.complex128):
        _case(gpu, n, 3, dtype, ff.FORWARD, inplace=True)
        _case(gpu, n, 3, dtype, ff.INVERSE, inplace=True)


def test_config_4096x256_forward_inverse(gpu):
    """C2C FP32 N=4096 batch 256, N(0,1) inputs: forward vs FP64 oracle and inverse/N roundtrip <= 2.5e-5."""
    n, batch = 4096, 256
    x = complex_input(0x4096, (batch, n), np.complex64, kind="normal")
    y = run_c2c(gpu, x, ff.FORWARD, n, batch)
    z = run_c2c(gpu, y, ff.INVERSE, n, batch)
    e_fwd, _ = ora.rel_l2(y, ora.fft(x.astype(np.complex128)))
    e_rt, _ = ora.rel_l2(z / n, x)
    assert e_fwd <= 2.5e-5 and e_rt <= 2.5e-5, (e_fwd, e_rt)


@pytest.mark.parametrize("n", [10007, 10935])
def test_config_arbitrary_lengths(gpu, n):
    """C2C FP32 and Z2Z FP64 at N=10007 (prime) and 10935 = 3^7*5, batch 64, N(0,1)."""
    for dtype, t in ((np.complex64, ff.Type.C2C), (np.complex128, ff.Type.Z2Z)):
        x = complex_input(0xC0FFEE + n, (64, n), dtype, kind="normal")
        y = run_c2c(gpu, x, ff.FORWARD, n, 64)
        check_accuracy(y, ora.fft(x.astype(np.complex128)), t, n, row_len=n, what="config")


def test_large_four_step(gpu):
    """C2C FP32 N=2^20 batch 16 and N=2^24 batch 1 (multi-pass Stockham)."""
    for n, batch in ((1 << 20, 16), (1 << 24, 1)):
        x = complex_input(0xB16 + n, (batch, n), np.complex64, kind="normal")
        y = run_c2c(gpu, x, ff.FORWARD, n, batch)
        check_accuracy(y, ora.fft(x.astype(np.complex128)), ff.Type.C2C, n, row_len=n, what="large")


# Lengths whose factors include the primes 17 and 19, which have radix
# codelets of their own (the reference's kSupportedRadices,
# src/utils/include/flagfft/base.hpp:47): single-pass leaves, multi-pass
# chains and a Rader convolution of length 102 = 17 * 6, never Bluestein.
RADIX_17_19 = [17, 19, 289, 361, 4352, 5168, 6859, 9728, 83521, 103, 2 * 17 * 19 * 1024]


@pytest.mark.parametrize("dtype", [np.complex64, np.complex128], ids=["c2c", "z2z"])
@pytest.mark.parametrize("n", RADIX_17_19)
def test_radix_17_and_19_lengths(gpu, n, dtype):
    t = ff.Type.Z2Z if dtype == np.complex128 else ff.Type.C2C
    r, text = ff.describe_host(1, [n], t, 3)
    assert r == ff.Result.SUCCESS and "Bluestein(" not in text, text
    for direction in (ff.FORWARD, ff.INVERSE):
        _case(gpu, n, 3, dtype, direction)


@pytest.mark.parametrize("n", [2 * 17, 2 * 361, 4352, 2 * 6859])
def test_radix_17_and_19_real_roundtrip(gpu, n):
    """R2C against the oracle, then C2R back to n x the input, in FP32 and FP64."""
    from .helpers import dev_tensor, host, real_input, stream_of, sync

    for dbl in (False, True):
        rdt = np.float64 if dbl else np.float32
        fwd, inv = (ff.Type.D2Z, ff.Type.Z2D) if dbl else (ff.Type.R2C, ff.Type.C2R)
        x = real_input(ora.accuracy_seed(int(fwd), n, 3, 0), (3, n), rdt)
        xd = dev_tensor(x, gpu)
        import torch

        yd = torch.empty((3, n // 2 + 1), dtype=torch.complex128 if dbl else torch.complex64, device=gpu)
        zd = torch.empty_like(xd)
        with ff.Plan.one_d(n, fwd, 3) as pf, ff.Plan.one_d(n, inv, 3) as pi:
            pf.set_stream(stream_of(gpu))
            pi.set_stream(stream_of(gpu))
            (pf.d2z if dbl else pf.r2c)(xd.data_ptr(), yd.data_ptr())
            (pi.z2d if dbl else pi.c2r)(yd.data_ptr(), zd.data_ptr())
            sync(gpu)
        check_accuracy(host(yd), ora.rfft(x.astype(np.float64)), fwd, n, row_len=n // 2 + 1, what="r2c")
        check_accuracy(host(zd), n * x.astype(np.float64), inv, n, row_len=n, factor=2.0, what="roundtrip")


# 5- and 10-smooth lengths run as chains of specialised LDS passes of 25,
# 100, 125 and 1000 points (k_stage [5x5], [10x10], [5x5x5], [10x10x10]).
# (400 = [10, 10, 4] and 120 = [15, 8] passes have no specialised instance;
# 10000 FP32 and 6000 are single-pass leaves with the sequence in LDS)
SMOOTH5 = [(10000, 3, "k_leaf<fp32,20,0>[10x10x10x10]", "[10x10]", True), (78125, 2, "[5x5]", "[5x5]", True),
           (100000, 1, "[10x10x10]", "[10x10x10]", True), (15625, 4, "[5x5x5]", "[5x5x5]", True),
           (1000000, 1, "[10x10x10]", "[10x10x10]", True), (250000, 1, "[5x5]", "[5x5]", True),
           (50000, 2, "[5x5x5]", "[5x5x5]", False), (12000, 2, "[10x10]", "[10x10]", False),
           (6000, 4, "k_leaf<fp32,16,0>[15x10x10x4]", "k_leaf<fp64,16,0>[15x10x10x4]", True)]


@pytest.mark.parametrize("dtype", [np.complex64, np.complex128], ids=["c2c", "z2z"])
@pytest.mark.parametrize("n,batch,inst32,inst64,all_specialised", SMOOTH5, ids=[f"{c[0]}" for c in SMOOTH5])
def test_five_smooth_sizes(gpu, n, batch, inst32, inst64, all_specialised, dtype):
    dbl = dtype == np.complex128
    t = ff.Type.Z2Z if dbl else ff.Type.C2C
    with ff.Plan.one_d(n, t, batch) as p:
        desc = p.description()
    inst = inst64 if dbl else inst32
    assert inst in desc and ("k_pass" not in desc) == all_specialised, desc
    for direction in (ff.FORWARD, ff.INVERSE):
        _case(gpu, n, batch, dtype, direction)


# 3, 5, 7, 9 and 15 times a power of two and 1000 times one: specialised
# single-pass leaves (FP64 lengths past 4096 whose sequence fits LDS as well).
SMOOTH_LEAVES = [384, 768, 1536, 3072, 6144, 640, 1280, 2560, 5120, 960, 1920, 3840, 1200, 2400, 4800, 3200, 6400,
                 1152, 2304, 4608, 896, 1792, 3584, 7168, 1440, 2880, 5760, 2000, 4000, 8000]


@pytest.mark.parametrize("dtype", [np.complex64, np.complex128], ids=["c2c", "z2z"])
@pytest.mark.parametrize("n", SMOOTH_LEAVES)
def test_smooth_leaf_sizes(gpu, n, dtype):
    t = ff.Type.Z2Z if dtype == np.complex128 else ff.Type.C2C
    with ff.Plan.one_d(n, t, 5) as p:
        desc = p.description()
    assert "k_leaf<" in desc and "k_pass" not in desc, desc
    for batch in (1, 5):
        for direction in (ff.FORWARD, ff.INVERSE):
            _case(gpu, n, batch, dtype, direction)
