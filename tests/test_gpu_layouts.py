"""Advanced layouts through flagfftxPlanMany (GPU): element strides, embeds and
distances for C2C / Z2Z of rank 1 and 2, and padded rows for real rank 2,
against the oracle on the gathered logical data.  Every position of the output buffer outside the output layout
must keep its sentinel value (no write outside the layout).
Bounds: FP32 2e-6 log2 N, FP64 1e-14 log2 N relative L2 per transform.
"""
import numpy as np
import pytest

from flagfft_amd import flagfft as ff
from oracle import oracle as ora

from .helpers import check_accuracy, complex_input, dev_tensor, host, real_input, stream_of, sync

pytestmark = pytest.mark.gpu

SENTINEL = 12345.0 - 678.0j


def _index(shape, emb, stride, dist, batch):
    """Flat element index of every logical (b, x, y) (rank 2) or (b, y) (rank 1)."""
    if len(shape) == 1:
        b, y = np.meshgrid(np.arange(batch), np.arange(shape[0]), indexing="ij")
        return b * dist + y * stride
    b, x, y = np.meshgrid(np.arange(batch), np.arange(shape[0]), np.arange(shape[1]), indexing="ij")
    return b * dist + (x * emb + y) * stride


def _run(gpu, shape, dbl, batch, istride, iemb, idist, ostride, oemb, odist, direction, in_place=False):
    t = ff.Type.Z2Z if dbl else ff.Type.C2C
    dtype = np.complex128 if dbl else np.complex64
    rank = len(shape)
    n = int(np.prod(shape))
    x = complex_input(n * 7 + istride + batch, (batch,) + tuple(shape), dtype)
    iidx = _index(shape, iemb, istride, idist, batch)
    oidx = _index(shape, oemb, ostride, odist, batch)
    size = max(int(iidx.max()), int(oidx.max())) + 1 + 5
    ibuf = np.full(size, SENTINEL, dtype)
    ibuf[iidx] = x
    obuf = ibuf if in_place else np.full(size, SENTINEL, dtype)
    inembed = [shape[0], iemb] if rank == 2 else [iemb]
    onembed = [shape[0], oemb] if rank == 2 else [oemb]
    with ff.Plan.many_ext(rank, list(shape), inembed, istride, idist, onembed, ostride, odist, t, batch) as p:
        desc = p.description()
        p.set_stream(stream_of(gpu))
        di = dev_tensor(ibuf, gpu)
        do = di if in_place else dev_tensor(obuf, gpu)
        (p.z2z if dbl else p.c2c)(di.data_ptr(), do.data_ptr(), direction)
        sync(gpu)
        out = host(do)
    ref = ora.fft(x.astype(np.complex128), sign=direction) if rank == 1 else ora.fft2(x.astype(np.complex128),
                                                                                       sign=direction)
    check_accuracy(out[oidx], ref, t, n, row_len=n, factor=2.0,
                   what=f"{shape} istride={istride} iemb={iemb} idist={idist} ostride={ostride} oemb={oemb} "
                        f"odist={odist} dir={direction} in_place={in_place}")
    outside = np.ones(size, bool)
    outside[oidx.reshape(-1)] = False
    if not in_place:
        assert np.all(out[outside] == SENTINEL), "write outside the output layout"
    return desc


ONE_D = [64, 1000, 4096, 97, 10007, 65536]


@pytest.mark.parametrize("dbl", [False, True], ids=["c2c", "z2z"])
@pytest.mark.parametrize("n", ONE_D)
def test_strided_1d(gpu, n, dbl):
    desc = _run(gpu, (n,), dbl, 3, 2, n, 2 * n + 5, 3, n, 3 * n + 1, ff.FORWARD)
    assert "layout: istride=2" in desc
    _run(gpu, (n,), dbl, 2, 3, n, 3 * n + 2, 1, n, n + 9, ff.INVERSE)


@pytest.mark.parametrize("dbl", [False, True], ids=["c2c", "z2z"])
@pytest.mark.parametrize("n", [4096, 8192, 10935, 6000])
def test_padded_distances_1d(gpu, n, dbl):
    """Unit stride, padded batch distances: the specialised kernels (split leaves included) serve them."""
    _run(gpu, (n,), dbl, 3, 1, n, n + 7, 1, n, n + 3, ff.FORWARD)
    _run(gpu, (n,), dbl, 3, 1, n, n + 7, 1, n, n + 7, ff.INVERSE, in_place=True)


TWO_D = [(64, 48), (23, 16), (128, 100), (16, 997), (4096, 8)]


@pytest.mark.parametrize("dbl", [False, True], ids=["c2c", "z2z"])
@pytest.mark.parametrize("shape", TWO_D, ids=lambda s: f"{s[0]}x{s[1]}")
def test_strided_2d(gpu, shape, dbl):
    n0, n1 = shape
    _run(gpu, shape, dbl, 2, 1, n1 + 3, n0 * (n1 + 3) + 4, 2, n1 + 5, 2 * n0 * (n1 + 5) + 1, ff.FORWARD)
    _run(gpu, shape, dbl, 2, 2, n1, 2 * n0 * n1 + 3, 1, n1 + 1, n0 * (n1 + 1), ff.INVERSE)
    _run(gpu, shape, dbl, 2, 2, n1 + 2, 2 * n0 * (n1 + 2) + 6, 2, n1 + 2, 2 * n0 * (n1 + 2) + 6, ff.FORWARD,
         in_place=True)


def test_reference_layouts_behave_as_plan_many(gpu):
    """Layouts flagfftPlanMany accepts give the same plan through the extension."""
    for args in ((1, [4096], None, 1, 4096, None, 1, 4096, ff.Type.C2C, 4),
                 (2, [64, 32], None, 1, 2048, None, 1, 2048, ff.Type.Z2Z, 2)):
        with ff.Plan.many(*args) as a, ff.Plan.many_ext(*args) as b:
            assert a.description() == b.description()


REAL_2D = [(64, 48), (23, 16), (128, 100), (16, 997), (32, 15), (4096, 64)]


@pytest.mark.parametrize("dbl", [False, True], ids=["fp32", "fp64"])
@pytest.mark.parametrize("shape", REAL_2D, ids=lambda s: f"{s[0]}x{s[1]}")
def test_real_2d_padded_rows(gpu, shape, dbl):
    """Real 2D with padded rows through flagfftxPlanMany: the in-place layout of
    2 (n1/2+1) reals per row (R2C then C2R in one buffer) and out-of-place rows
    with extra pitch on both sides; batch 2, n0 rows per batch entry."""
    import torch

    n0, n1 = shape
    h = n1 // 2 + 1
    n = n0 * n1
    batch = 2
    tf, ti = (ff.Type.D2Z, ff.Type.Z2D) if dbl else (ff.Type.R2C, ff.Type.C2R)
    rdt, cdt = (np.float64, np.complex128) if dbl else (np.float32, np.complex64)
    tdt = torch.float64 if dbl else torch.float32
    x = real_input(n + 17, (batch, n0, n1), rdt, kind="normal")
    ref = ora.rfft2(x.astype(np.float64))
    # in place: rows of 2h reals / h complex
    buf = torch.zeros(batch * n0 * 2 * h, dtype=tdt, device=gpu)
    buf.view(batch, n0, 2 * h)[:, :, :n1] = torch.from_numpy(x).to(gpu)
    pad_r, pad_c = 2 * h, h
    with ff.Plan.many_ext(2, [n0, n1], [n0, pad_r], 1, n0 * pad_r, [n0, pad_c], 1, n0 * pad_c, tf, batch) as pf, \
            ff.Plan.many_ext(2, [n0, n1], [n0, pad_c], 1, n0 * pad_c, [n0, pad_r], 1, n0 * pad_r, ti, batch) as pi:
        assert "layout:" in pf.description()
        pf.set_stream(stream_of(gpu))
        pi.set_stream(stream_of(gpu))
        (pf.d2z if dbl else pf.r2c)(buf.data_ptr(), buf.data_ptr())
        sync(gpu)
        X = host(buf).view(cdt).reshape(batch, n0, h)
        check_accuracy(X, ref, tf, n, row_len=n0 * h, what="in-place padded R2C")
        (pi.z2d if dbl else pi.c2r)(buf.data_ptr(), buf.data_ptr())
        sync(gpu)
        z = host(buf).reshape(batch, n0, 2 * h)[:, :, :n1]
        check_accuracy(z, x.astype(np.float64) * n, ti, n, row_len=n, factor=2.0, what="in-place padded roundtrip")
    # out of place: input rows n1 + 6 reals, output rows h + 3 complex, and back
    pin, pout = n1 + 6, h + 3
    xin = np.full((batch, n0, pin), 7.0, rdt)
    xin[:, :, :n1] = x
    xd = dev_tensor(xin, gpu)
    Yd = dev_tensor(np.full((batch, n0, pout), SENTINEL, cdt), gpu)
    zd = dev_tensor(np.full((batch, n0, pin), -3.0, rdt), gpu)
    with ff.Plan.many_ext(2, [n0, n1], [n0, pin], 1, n0 * pin, [n0, pout], 1, n0 * pout, tf, batch) as pf, \
            ff.Plan.many_ext(2, [n0, n1], [n0, pout], 1, n0 * pout, [n0, pin], 1, n0 * pin, ti, batch) as pi:
        pf.set_stream(stream_of(gpu))
        pi.set_stream(stream_of(gpu))
        (pf.d2z if dbl else pf.r2c)(xd.data_ptr(), Yd.data_ptr())
        (pi.z2d if dbl else pi.c2r)(Yd.data_ptr(), zd.data_ptr())
        sync(gpu)
    Y = host(Yd)
    check_accuracy(Y[:, :, :h], ref, tf, n, row_len=n0 * h, what="padded R2C")
    assert np.all(Y[:, :, h:] == SENTINEL), "R2C wrote into the row padding"
    z = host(zd)
    check_accuracy(z[:, :, :n1], x.astype(np.float64) * n, ti, n, row_len=n, factor=2.0, what="padded roundtrip")
    assert np.all(z[:, :, n1:] == -3.0), "C2R wrote into the row padding"
