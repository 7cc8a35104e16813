"""Rank-3 transforms (flagfftxPlan3d) against the oracle (GPU).

The oracle composes its 1D / 2D transforms over the three axes (pinned
against a library routine and a closed form in test_oracle.py).  Shapes mix
single-pass axes, Rader (23, 67) and Bluestein-sized axes, odd real lengths,
long columns (multi-pass column chains) and batches.
Bounds: FP32 2e-6 log2 N, FP64 1e-14 log2 N relative L2 (N = n0 n1 n2).
"""
import numpy as np
import pytest

from flagfft_amd import flagfft as ff
from oracle import oracle as ora

from .helpers import check_accuracy, complex_input, dev_tensor, host, real_input, stream_of, sync

pytestmark = pytest.mark.gpu

SHAPES = [(16, 16, 16), (8, 12, 10), (64, 32, 16), (23, 9, 5), (5, 7, 67), (3, 17, 64), (128, 128, 128),
          (4096, 4, 3)]


@pytest.mark.parametrize("dbl", [False, True], ids=["c2c", "z2z"])
@pytest.mark.parametrize("shape", SHAPES, ids=lambda s: "x".join(map(str, s)))
def test_c2c_3d(gpu, shape, dbl):
    n0, n1, n2 = shape
    n = n0 * n1 * n2
    t = ff.Type.Z2Z if dbl else ff.Type.C2C
    for batch in ((1, 2) if n <= 1 << 16 else (1,)):
        x = complex_input(n + batch, (batch, n0, n1, n2), np.complex128 if dbl else np.complex64)
        xd = dev_tensor(x, gpu)
        yd = xd.clone().zero_()
        zd = xd.clone().zero_()
        with ff.Plan.three_d(n0, n1, n2, t, batch) as p:
            assert "ThreeDim(" in p.description()
            p.set_stream(stream_of(gpu))
            ex = p.z2z if dbl else p.c2c
            ex(xd.data_ptr(), yd.data_ptr(), ff.FORWARD)
            ex(yd.data_ptr(), zd.data_ptr(), ff.INVERSE)
            sync(gpu)
        check_accuracy(host(yd), ora.fft3(x.astype(np.complex128)), t, n, row_len=n, factor=2.0,
                       what=f"3D fwd batch={batch}")
        check_accuracy(host(zd), x.astype(np.complex128) * n, t, n, row_len=n, factor=4.0, what="3D roundtrip")


REAL_SHAPES = [(16, 8, 6), (12, 10, 9), (32, 16, 20), (7, 5, 64), (23, 4, 15), (128, 128, 128), (4096, 2, 6)]


@pytest.mark.parametrize("dbl", [False, True], ids=["fp32", "fp64"])
@pytest.mark.parametrize("shape", REAL_SHAPES, ids=lambda s: "x".join(map(str, s)))
def test_real_3d(gpu, shape, dbl):
    n0, n1, n2 = shape
    n = n0 * n1 * n2
    h = n2 // 2 + 1
    tf, ti = (ff.Type.D2Z, ff.Type.Z2D) if dbl else (ff.Type.R2C, ff.Type.C2R)
    rdt, cdt = (np.float64, np.complex128) if dbl else (np.float32, np.complex64)
    for batch in ((1, 2) if n <= 1 << 16 else (1,)):
        x = real_input(n + 3 * batch, (batch, n0, n1, n2), rdt, kind="normal")
        xd = dev_tensor(x, gpu)
        Xd = dev_tensor(np.zeros((batch, n0, n1, h), cdt), gpu)
        zd = xd.clone().zero_()
        with ff.Plan.three_d(n0, n1, n2, tf, batch) as pf, ff.Plan.three_d(n0, n1, n2, ti, batch) as pi:
            pf.set_stream(stream_of(gpu))
            pi.set_stream(stream_of(gpu))
            (pf.d2z if dbl else pf.r2c)(xd.data_ptr(), Xd.data_ptr())
            (pi.z2d if dbl else pi.c2r)(Xd.data_ptr(), zd.data_ptr())
            sync(gpu)
        X = host(Xd)
        check_accuracy(X, ora.rfft3(x.astype(np.float64)), tf, n, row_len=n0 * n1 * h, what=f"R2C batch={batch}")
        check_accuracy(host(zd), x.astype(np.float64) * n, ti, n, row_len=n, factor=2.0, what="roundtrip")
        # C2R of a spectrum that is not Hermitian on the DC / Nyquist planes:
        # the oracle's definition (Im of those bins of the last axis ignored)
        Y = complex_input(n + 7, (batch, n0, n1, h), cdt)
        Yd = dev_tensor(Y, gpu)
        yd = xd.clone().zero_()
        with ff.Plan.three_d(n0, n1, n2, ti, batch) as pi:
            pi.set_stream(stream_of(gpu))
            (pi.z2d if dbl else pi.c2r)(Yd.data_ptr(), yd.data_ptr())
            sync(gpu)
        assert np.array_equal(host(Yd), Y), "C2R must not modify its input"
        check_accuracy(host(yd), ora.irfft3(Y.astype(np.complex128), n2), ti, n, row_len=n, factor=2.0,
                       what="C2R vs oracle")


@pytest.mark.parametrize("shape", [(16, 8, 6), (9, 10, 13), (64, 64, 64)], ids=lambda s: "x".join(map(str, s)))
def test_real_3d_in_place(gpu, shape):
    """R2C then C2R with the input and output in one buffer of max(in, out) bytes."""
    import torch

    n0, n1, n2 = shape
    n = n0 * n1 * n2
    h = n2 // 2 + 1
    x = real_input(n + 11, (1, n0, n1, n2), np.float32, kind="normal")
    buf = torch.zeros(2 * n0 * n1 * h, dtype=torch.float32, device=gpu)
    buf[:n] = torch.from_numpy(x.reshape(-1)).to(gpu)
    with ff.Plan.three_d(n0, n1, n2, ff.Type.R2C) as pf, ff.Plan.three_d(n0, n1, n2, ff.Type.C2R) as pi:
        pf.set_stream(stream_of(gpu))
        pi.set_stream(stream_of(gpu))
        pf.r2c(buf.data_ptr(), buf.data_ptr())
        sync(gpu)
        X = host(buf).view(np.complex64).reshape(1, n0, n1, h)
        pi.c2r(buf.data_ptr(), buf.data_ptr())
        sync(gpu)
        z = host(buf)[:n].reshape(1, n0, n1, n2)
    check_accuracy(X, ora.rfft3(x.astype(np.float64)), ff.Type.R2C, n, row_len=n0 * n1 * h, what="in-place R2C")
    check_accuracy(z, x.astype(np.float64) * n, ff.Type.C2R, n, row_len=n, factor=2.0, what="in-place roundtrip")


@pytest.mark.parametrize("shape", [(1, 1, 1), (1, 8, 1), (7, 1, 5), (2, 1, 1), (1, 1, 16), (3, 2, 2), (1, 5, 2)],
                         ids=lambda s: "x".join(map(str, s)))
def test_degenerate_axes(gpu, shape):
    """Axes of length 1 and 2 (length-1 transforms are the identity) for C2C, Z2Z, R2C and C2R."""
    n0, n1, n2 = shape
    n = n0 * n1 * n2
    h = n2 // 2 + 1
    x = complex_input(n + 101, (1, n0, n1, n2), np.complex128)
    for t, dtype in ((ff.Type.C2C, np.complex64), (ff.Type.Z2Z, np.complex128)):
        xd = dev_tensor(x.astype(dtype), gpu)
        yd = xd.clone().zero_()
        with ff.Plan.three_d(n0, n1, n2, t) as p:
            p.set_stream(stream_of(gpu))
            (p.z2z if t == ff.Type.Z2Z else p.c2c)(xd.data_ptr(), yd.data_ptr(), ff.FORWARD)
            sync(gpu)
        check_accuracy(host(yd), ora.fft3(x.astype(dtype).astype(np.complex128)), t, max(n, 2), row_len=n,
                       factor=2.0, what="degenerate C2C")
    xr = x.real.astype(np.float32)
    xd = dev_tensor(xr, gpu)
    Xd = dev_tensor(np.zeros((1, n0, n1, h), np.complex64), gpu)
    zd = xd.clone().zero_()
    with ff.Plan.three_d(n0, n1, n2, ff.Type.R2C) as pf, ff.Plan.three_d(n0, n1, n2, ff.Type.C2R) as pi:
        pf.set_stream(stream_of(gpu))
        pi.set_stream(stream_of(gpu))
        pf.r2c(xd.data_ptr(), Xd.data_ptr())
        pi.c2r(Xd.data_ptr(), zd.data_ptr())
        sync(gpu)
    check_accuracy(host(Xd), ora.rfft3(xr.astype(np.float64)), ff.Type.R2C, max(n, 2), row_len=n0 * n1 * h,
                   factor=2.0, what="degenerate R2C")
    check_accuracy(host(zd), xr.astype(np.float64) * n, ff.Type.C2R, max(n, 2), row_len=n, factor=2.0,
                   what="degenerate roundtrip")


def test_c2c_3d_in_place_and_graph(gpu):
    n0, n1, n2 = 32, 24, 40
    n = n0 * n1 * n2
    x = complex_input(n, (2, n0, n1, n2), np.complex64)
    ref = ora.fft3(x.astype(np.complex128))
    with ff.Plan.three_d(n0, n1, n2, ff.Type.C2C, 2) as p:
        p.set_stream(stream_of(gpu))
        xd = dev_tensor(x, gpu)
        yd = xd.clone().zero_()
        p.c2c(xd.data_ptr(), yd.data_ptr(), ff.FORWARD)
        zd = xd.clone()
        p.c2c(zd.data_ptr(), zd.data_ptr(), ff.FORWARD)
        sync(gpu)
        assert np.array_equal(host(zd), host(yd))
        p.set_graph_mode(True)
        for rep in range(3):
            gd = xd.clone().zero_()
            p.c2c(xd.data_ptr(), gd.data_ptr(), ff.FORWARD)
            sync(gpu)
            assert np.array_equal(host(gd), host(yd)), f"graph replay {rep}"
    check_accuracy(host(yd), ref, ff.Type.C2C, n, row_len=n, factor=2.0, what="3D")


def test_3d_exec_contract(gpu):
    """Exec on a rank-3 plan follows the reference's Exec contract (src/exec/c_api_exec.cpp:9-21):
    wrong type -> INVALID_TYPE, null pointer -> INVALID_VALUE, bad direction -> INVALID_VALUE,
    destroyed handle -> INVALID_PLAN."""
    import torch

    R = ff.Result
    x = torch.zeros(2 * 4 * 6 * 8, dtype=torch.complex64, device=gpu)
    p = ff.Plan.three_d(4, 6, 8, ff.Type.C2C, 2)
    h = p.handle
    assert ff.exec_c2c(h, x.data_ptr(), x.data_ptr(), 7) == R.INVALID_VALUE
    assert ff.exec_c2c(h, 0, x.data_ptr(), ff.FORWARD) == R.INVALID_VALUE
    assert ff.exec_z2z(h, x.data_ptr(), x.data_ptr(), ff.FORWARD) == R.INVALID_TYPE
    assert ff.exec_c2c(h, x.data_ptr(), x.data_ptr(), ff.FORWARD) == R.SUCCESS
    sync(gpu)
    p.destroy()
    assert ff.exec_c2c(h, x.data_ptr(), x.data_ptr(), ff.FORWARD) == R.INVALID_PLAN
