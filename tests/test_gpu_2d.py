"""2D transforms against the oracle (GPU).

Mirrors ctest/test_2d_correctness.cpp: its sizes {16^2, 64^2, 128x256, 256x128,
256^2} and non-smooth {23^2, 997x16, 16x997}, batch {1, 4} via PlanMany,
forward/inverse roundtrip, R2C/C2R, plus the 4096^2 R2C/C2R and 2048^2 Z2Z configs.
"""
import numpy as np
import pytest

from flagfft_amd import flagfft as ff
from oracle import oracle as ora

from .helpers import check_accuracy, complex_input, dev_tensor, host, real_input, stream_of, sync

pytestmark = pytest.mark.gpu

SHAPES = [(16, 16), (64, 64), (128, 256), (256, 128), (256, 256), (23, 23), (997, 16), (16, 997), (67, 32)]


def _plan(n0, n1, t, batch):
    if batch == 1:
        return ff.Plan.two_d(n0, n1, t)
    real_fwd = t in (ff.Type.R2C, ff.Type.D2Z)
    real_inv = t in (ff.Type.C2R, ff.Type.Z2D)
    h = n1 // 2 + 1
    idist = n0 * h if real_inv else n0 * n1
    odist = n0 * h if real_fwd else n0 * n1
    return ff.Plan.many(2, [n0, n1], None, 1, idist, None, 1, odist, t, batch)


@pytest.mark.parametrize("dbl", [False, True], ids=["c2c", "z2z"])
@pytest.mark.parametrize("shape", SHAPES, ids=lambda s: f"{s[0]}x{s[1]}")
def test_c2c_2d(gpu, shape, dbl):
    n0, n1 = shape
    t = ff.Type.Z2Z if dbl else ff.Type.C2C
    for batch in (1, 4):
        x = complex_input(n0 * 7919 + n1 + batch, (batch, n0, n1), np.complex128 if dbl else np.complex64)
        xd = dev_tensor(x, gpu)
        yd = xd.clone().zero_()
        zd = xd.clone().zero_()
        with _plan(n0, n1, t, batch) as p:
            p.set_stream(stream_of(gpu))
            ex = p.z2z if dbl else p.c2c
            ex(xd.data_ptr(), yd.data_ptr(), ff.FORWARD)
            ex(yd.data_ptr(), zd.data_ptr(), ff.INVERSE)
            sync(gpu)
        ref = ora.fft2(x.astype(np.complex128))
        n = n0 * n1
        check_accuracy(host(yd), ref, t, n, row_len=n, factor=2.0, what=f"2D fwd batch={batch}")
        check_accuracy(host(zd), x.astype(np.complex128) * n, t, n, row_len=n, factor=4.0, what="2D roundtrip")



# Column passes split into super-stages (the explicit split applies to the
# column length only): specialised column super-stages with inter-pass
# twiddles, mixed with run-time-radix passes where a super-radix has no
# specialised instance (16).
CHAINS = [((4096, 48), False, [64, 64]), ((8192, 20), False, [128, 64]), ((2048, 40), True, [16, 128]),
          ((1024, 36), True, [64, 16]), ((4096, 24), True, [16, 16, 16])]


@pytest.mark.parametrize("shape,dbl,passes", CHAINS, ids=[f"{s[0]}x{s[1]}-{'z2z' if d else 'c2c'}-{'x'.join(map(str, p))}"
                                                          for s, d, p in CHAINS])
def test_column_chains(gpu, shape, dbl, passes):
    n0, n1 = shape
    t = ff.Type.Z2Z if dbl else ff.Type.C2C
    batch = 2
    x = complex_input(n0 + 31 * n1, (batch, n0, n1), np.complex128 if dbl else np.complex64)
    xd = dev_tensor(x, gpu)
    yd = xd.clone().zero_()
    zd = xd.clone().zero_()
    ff.set_tuning(passes=passes)
    try:
        p = _plan(n0, n1, t, batch)
    finally:
        ff.set_tuning()
    with p:
        assert "MultiPassStockham(n=%d" % n0 in p.description()
        p.set_stream(stream_of(gpu))
        ex = p.z2z if dbl else p.c2c
        ex(xd.data_ptr(), yd.data_ptr(), ff.FORWARD)
        ex(yd.data_ptr(), zd.data_ptr(), ff.INVERSE)
        sync(gpu)
    n = n0 * n1
    check_accuracy(host(yd), ora.fft2(x.astype(np.complex128)), t, n, row_len=n, factor=2.0, what="2D fwd")
    check_accuracy(host(zd), x.astype(np.complex128) * n, t, n, row_len=n, factor=4.0, what="2D roundtrip")


@pytest.mark.parametrize("dbl", [False, True], ids=["fp32", "fp64"])
@pytest.mark.parametrize("shape", [(64, 64), (67, 32), (256, 130), (4096, 4096)], ids=lambda s: f"{s[0]}x{s[1]}")
def test_real_2d_in_place(gpu, shape, dbl):
    """R2C then C2R with the input and output in one buffer of max(in, out) bytes."""
    import torch

    n0, n1 = shape
    h = n1 // 2 + 1
    tf, ti = (ff.Type.D2Z, ff.Type.Z2D) if dbl else (ff.Type.R2C, ff.Type.C2R)
    rdt = np.float64 if dbl else np.float32
    x = real_input(n0 * 3 + n1, (1, n0, n1), rdt, kind="normal")
    buf = torch.zeros(2 * n0 * h, dtype=torch.float64 if dbl else torch.float32, device=gpu)
    buf[: n0 * n1] = torch.from_numpy(x.reshape(-1)).to(gpu)
    with ff.Plan.two_d(n0, n1, tf) as pf, ff.Plan.two_d(n0, n1, ti) as pi:
        pf.set_stream(stream_of(gpu))
        pi.set_stream(stream_of(gpu))
        (pf.d2z if dbl else pf.r2c)(buf.data_ptr(), buf.data_ptr())
        sync(gpu)
        X = host(buf).view(np.complex128 if dbl else np.complex64).reshape(1, n0, h)
        (pi.z2d if dbl else pi.c2r)(buf.data_ptr(), buf.data_ptr())
        sync(gpu)
        z = host(buf)[: n0 * n1].reshape(1, n0, n1)
    n = n0 * n1
    check_accuracy(X, ora.rfft2(x.astype(np.float64)), tf, n, row_len=n0 * h, what="in-place R2C")
    check_accuracy(z, x.astype(np.float64) * n, ti, n, row_len=n, factor=2.0, what="in-place roundtrip")


@pytest.mark.parametrize("dbl", [False, True], ids=["fp32", "fp64"])
@pytest.mark.parametrize("shape", SHAPES + [(33, 15)], ids=lambda s: f"{s[0]}x{s[1]}")
def test_real_2d(gpu, shape, dbl):
    n0, n1 = shape
    h = n1 // 2 + 1
    tf = ff.Type.D2Z if dbl else ff.Type.R2C
    ti = ff.Type.Z2D if dbl else ff.Type.C2R
    for batch in (1, 2):
        x = real_input(n0 + 31 * n1 + batch, (batch, n0, n1), np.float64 if dbl else np.float32)
        xd = dev_tensor(x, gpu)
        Xd = dev_tensor(np.zeros((batch, n0, h), np.complex128 if dbl else np.complex64), gpu)
        zd = xd.clone().zero_()
        with _plan(n0, n1, tf, batch) as pf, _plan(n0, n1, ti, batch) as pi:
            pf.set_stream(stream_of(gpu))
            pi.set_stream(stream_of(gpu))
            (pf.d2z if dbl else pf.r2c)(xd.data_ptr(), Xd.data_ptr())
            (pi.z2d if dbl else pi.c2r)(Xd.data_ptr(), zd.data_ptr())
            sync(gpu)
        n = n0 * n1
        check_accuracy(host(Xd), ora.rfft2(x.astype(np.float64)), tf, n, row_len=n0 * h, factor=2.0,
                       what=f"2D R2C batch={batch}")
        check_accuracy(host(zd), x.astype(np.float64) * n, ti, n, row_len=n, factor=4.0, what="2D roundtrip")


def test_c2r_2d_against_oracle(gpu):
    """2D C2R of a non-Hermitian input: the row pass drops Im of each row's DC/Nyquist bins."""
    n0, n1 = 24, 20
    h = n1 // 2 + 1
    X = complex_input(99, (1, n0, h), np.complex64)
    Xd = dev_tensor(X, gpu)
    yd = dev_tensor(np.zeros((1, n0, n1), np.float32), gpu)
    with ff.Plan.two_d(n0, n1, ff.Type.C2R) as p:
        p.set_stream(stream_of(gpu))
        p.c2r(Xd.data_ptr(), yd.data_ptr())
        sync(gpu)
    check_accuracy(host(yd), ora.irfft2(X.astype(np.complex128), n1), ff.Type.C2R, n0 * n1, row_len=n0 * n1,
                   factor=2.0)


def test_config_2d(gpu):
    """2D R2C/C2R FP32 4096x4096 (N(0,1)) and 2D Z2Z FP64 2048x2048."""
    n0 = n1 = 4096
    h = n1 // 2 + 1
    x = real_input(0x2D, (1, n0, n1), np.float32, kind="normal")
    xd = dev_tensor(x, gpu)
    Xd = dev_tensor(np.zeros((1, n0, h), np.complex64), gpu)
    zd = xd.clone().zero_()
    with ff.Plan.two_d(n0, n1, ff.Type.R2C) as pf, ff.Plan.two_d(n0, n1, ff.Type.C2R) as pi:
        pf.set_stream(stream_of(gpu))
        pi.set_stream(stream_of(gpu))
        pf.r2c(xd.data_ptr(), Xd.data_ptr())
        pi.c2r(Xd.data_ptr(), zd.data_ptr())
        sync(gpu)
    n = n0 * n1
    check_accuracy(host(Xd), ora.rfft2(x.astype(np.float64)), ff.Type.R2C, n, row_len=n0 * h, what="4096^2 R2C")
    check_accuracy(host(zd), x.astype(np.float64) * n, ff.Type.C2R, n, row_len=n, factor=2.0, what="4096^2 rt")
    n0 = n1 = 2048
    x = complex_input(0x2E, (1, n0, n1), np.complex128, kind="normal")
    xd = dev_tensor(x, gpu)
    yd = xd.clone().zero_()
    with ff.Plan.two_d(n0, n1, ff.Type.Z2Z) as p:
        p.set_stream(stream_of(gpu))
        p.z2z(xd.data_ptr(), yd.data_ptr(), ff.FORWARD)
        sync(gpu)
    check_accuracy(host(yd), ora.fft2(x), ff.Type.Z2Z, n0 * n1, row_len=n0 * n1, what="2048^2 Z2Z")


# 4096-point columns run as one register-resident pass of 8 columns per
# workgroup (64 x 64 points in registers) when the tiles fill the CUs in
# whole rounds; with one transform the last tile may be partial (ncol not a
# multiple of 8: its idle lanes must neither read nor write), with several
# each transform needs whole tiles.  Other column counts take two 64-point
# passes.
@pytest.mark.parametrize("shape,batch,single", [((4096, 4093), 1, True), ((4096, 2045), 1, True),
                                                ((4096, 1024), 2, True), ((4096, 333), 1, False)],
                         ids=["4096x4093", "4096x2045", "4096x1024-b2", "4096x333"])
def test_4096_point_columns(gpu, shape, batch, single):
    n0, n1 = shape
    x = complex_input(n0 + 3 * n1 + batch, (batch, n0, n1), np.complex64)
    xd = dev_tensor(x, gpu)
    yd = xd.clone().zero_()
    with _plan(n0, n1, ff.Type.C2C, batch) as p:
        desc = p.description()
        p.set_stream(stream_of(gpu))
        p.c2c(xd.data_ptr(), yd.data_ptr(), ff.FORWARD)
        sync(gpu)
    assert ("k_col<fp32,8,64,64" in desc) == single, desc
    ref = ora.fft2(x.astype(np.complex128))
    n = n0 * n1
    check_accuracy(host(yd).reshape(batch, n), ref.reshape(batch, n), ff.Type.C2C, n, row_len=n, factor=2.0,
                   what=f"2D {n0}x{n1} batch {batch}")


# 8192-point columns have no single-pass instance that reads wide row
# segments (the LDS form would read one element per column): two passes of
# 128 and 64 points.
@pytest.mark.parametrize("shape,t", [((8192, 100), ff.Type.C2C), ((8192, 256), ff.Type.R2C),
                                     ((8192, 36), ff.Type.Z2Z), ((8192, 64), ff.Type.D2Z)],
                         ids=["c2c-8192x100", "r2c-8192x256", "z2z-8192x36", "d2z-8192x64"])
def test_8192_point_columns(gpu, shape, t):
    import torch

    n0, n1 = shape
    real = t in (ff.Type.R2C, ff.Type.D2Z)
    dbl = t in (ff.Type.Z2Z, ff.Type.D2Z)
    h = n1 // 2 + 1 if real else n1
    if real:
        x = real_input(n0 + n1, (1, n0, n1), np.float64 if dbl else np.float32)
        ref = ora.rfft2(x.astype(np.float64))
    else:
        x = complex_input(n0 + n1, (1, n0, n1), np.complex128 if dbl else np.complex64)
        ref = ora.fft2(x.astype(np.complex128))
    xd = dev_tensor(x, gpu)
    yd = torch.zeros((1, n0, h), dtype=torch.complex128 if dbl else torch.complex64, device=gpu)
    with ff.Plan.two_d(n0, n1, t) as p:
        desc = p.description()
        assert "super_radices=[128,64]" in desc, desc
        p.set_stream(stream_of(gpu))
        {ff.Type.C2C: lambda: p.c2c(xd.data_ptr(), yd.data_ptr(), ff.FORWARD),
         ff.Type.Z2Z: lambda: p.z2z(xd.data_ptr(), yd.data_ptr(), ff.FORWARD),
         ff.Type.R2C: lambda: p.r2c(xd.data_ptr(), yd.data_ptr()),
         ff.Type.D2Z: lambda: p.d2z(xd.data_ptr(), yd.data_ptr())}[t]()
        sync(gpu)
    n = n0 * n1
    check_accuracy(host(yd).reshape(1, -1), ref.reshape(1, -1), t, n, row_len=n0 * h, factor=2.0,
                   what=f"2D {n0}x{n1}")


# One real transform's n1 / 2 + 1 half-spectrum columns of 4096 points would
# leave one 8-column tile alone on the GPU: the DC and Nyquist columns (real on
# the real side) share unit 0 of a single register-resident pass instead
# (R2C: split after the transform; C2R: the Hermitian parts of both columns
# enter it); several transforms take two 64-point passes.
@pytest.mark.parametrize("shape,batch,paired", [((4096, 4096), 1, True), ((4096, 8192), 1, True),
                                                ((4096, 4096), 2, False)],
                         ids=["4096x4096", "4096x8192", "4096x4096-b2"])
def test_4096_point_columns_of_real_halves(gpu, shape, batch, paired):
    import torch

    n0, n1 = shape
    h = n1 // 2 + 1
    x = real_input(0x4096 + n1 + batch, (batch, n0, n1), np.float32, kind="normal")
    xd = dev_tensor(x, gpu)
    yd = torch.zeros((batch, n0, h), dtype=torch.complex64, device=gpu)
    zd = torch.zeros_like(xd)
    with _plan(n0, n1, ff.Type.R2C, batch) as pf, _plan(n0, n1, ff.Type.C2R, batch) as pi:
        df, di = pf.description(), pi.description()
        assert ("dc_nyquist_pair" in df and "dc-nyquist-unpack" in df) == paired, df
        assert ("dc_nyquist_pair" in di and "dc-nyquist-pack" in di) == paired, di
        assert ("super_radices=[64,64]" in df) != paired, df
        for p in (pf, pi):
            p.set_stream(stream_of(gpu))
        pf.r2c(xd.data_ptr(), yd.data_ptr())
        pi.c2r(yd.data_ptr(), zd.data_ptr())
        sync(gpu)
    n = n0 * n1
    ref = ora.rfft2(x.astype(np.float64))
    X = host(yd)
    check_accuracy(X.reshape(batch, -1), ref.reshape(batch, -1), ff.Type.R2C, n, row_len=n0 * h, factor=2.0,
                   what=f"2D R2C {n0}x{n1} batch {batch}")
    for c in (0, 1, h - 2, h - 1):  # the paired columns and their neighbours, each on its own
        check_accuracy(X[:, :, c], ref[:, :, c], ff.Type.R2C, n, row_len=n0, factor=2.0, what=f"column {c}")
    check_accuracy(host(zd).reshape(batch, -1), n * x.astype(np.float64).reshape(batch, -1), ff.Type.C2R, n,
                   row_len=n, factor=4.0, what=f"2D C2R roundtrip {n0}x{n1} batch {batch}")


def test_c2r_2d_paired_columns_of_non_hermitian_input(gpu):
    """C2R 4096 x 4096 of a non-Hermitian spectrum through the paired DC /
    Nyquist unit: the result is the reference's (column transforms of every
    half-spectrum column, then each row's C2R drops Im of its DC and Nyquist
    bins), which the paired unit reaches through the Hermitian parts of the two
    columns."""
    n0 = n1 = 4096
    h = n1 // 2 + 1
    X = complex_input(0xC2F, (1, n0, h), np.complex64, kind="normal")
    Xd = dev_tensor(X, gpu)
    yd = dev_tensor(np.zeros((1, n0, n1), np.float32), gpu)
    with ff.Plan.two_d(n0, n1, ff.Type.C2R) as p:
        assert "dc-nyquist-pack" in p.description()
        p.set_stream(stream_of(gpu))
        p.c2r(Xd.data_ptr(), yd.data_ptr())
        sync(gpu)
    check_accuracy(host(yd), ora.irfft2(X.astype(np.complex128), n1), ff.Type.C2R, n0 * n1, row_len=n0 * n1,
                   factor=2.0, what="non-Hermitian C2R 4096x4096")
