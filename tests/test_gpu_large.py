"""Jobs past 2^31 complex elements (GPU): the single-pass leaf (N = 4096,
batch 2^19 + 3) and the two-pass Stockham chain (N = 2^24, batch 129, with
its workspace of the same size) against the oracle on sampled sequences on
both sides of the 2^29-element (4 GiB) and 2^31-element boundaries, where a
32-bit byte or element offset would wrap.  About 35 GB (leaf) and 52 GB
(chain) of device memory; the FP64 Bluestein case about 103 GB.
Bounds: FP32 2e-6 log2 N, FP64 1e-14 log2 N relative L2 per sampled transform.
"""
import numpy as np
import pytest

from flagfft_amd import flagfft as ff
from oracle import oracle as ora

from .helpers import check_accuracy, host, stream_of, sync

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("n,batch", [(4096, (1 << 19) + 3), (1 << 24, 129)], ids=["leaf-4096", "chain-2^24"])
def test_past_2_31_elements(gpu, n, batch):
    import torch

    total = n * batch
    assert total > 1 << 31
    gen = torch.Generator(device=gpu)
    gen.manual_seed(n + batch)
    x = torch.randn(total, dtype=torch.complex64, device=gpu, generator=gen)
    y = torch.empty_like(x)
    try:
        with ff.Plan.one_d(n, ff.Type.C2C, batch) as p:
            p.set_stream(stream_of(gpu))
            p.c2c(x.data_ptr(), y.data_ptr(), ff.FORWARD)
            sync(gpu)
        last = batch - 1
        rows = sorted({0, ((1 << 29) // n) - 1, (1 << 29) // n, ((1 << 31) // n) - 1, (1 << 31) // n, last})
        for r in rows:
            xr = host(x[r * n:(r + 1) * n]).astype(np.complex128)
            yr = host(y[r * n:(r + 1) * n])
            check_accuracy(yr[None], ora.fft(xr[None]), ff.Type.C2C, n, row_len=n, what=f"N={n} sequence {r}")
    finally:
        del x, y
        torch.cuda.empty_cache()


def _rows(n, batch, elems_per_row):
    """Sampled rows: first, last, and either side of the 2^29- and 2^31-element marks."""
    rows = {0, batch - 1}
    for mark in (1 << 29, 1 << 31):
        r = mark // elems_per_row
        rows |= {r - 1, r} if r < batch else set()
    return sorted(rows)


@pytest.mark.parametrize("dbl", [False, True], ids=["c2c", "z2z"])
def test_bluestein_past_2_31_elements(gpu, dbl):
    """N = 10007 (Bluestein in LDS: the whole convolution in one workgroup for
    FP32, two halves with a parked first half for FP64), batch 214601."""
    import torch

    n, batch = 10007, 214601
    t = ff.Type.Z2Z if dbl else ff.Type.C2C
    assert n * batch > 1 << 31
    gen = torch.Generator(device=gpu)
    gen.manual_seed(n + dbl)
    x = torch.randn(n * batch, dtype=torch.complex128 if dbl else torch.complex64, device=gpu, generator=gen)
    y = torch.empty_like(x)
    try:
        with ff.Plan.one_d(n, t, batch) as p:
            assert "Bluestein" in p.description()
            p.set_stream(stream_of(gpu))
            (p.z2z if dbl else p.c2c)(x.data_ptr(), y.data_ptr(), ff.FORWARD)
            sync(gpu)
        for r in _rows(n, batch, n):
            xr = host(x[r * n:(r + 1) * n]).astype(np.complex128)
            check_accuracy(host(y[r * n:(r + 1) * n])[None], ora.fft(xr[None]), t, n, row_len=n,
                           what=f"N={n} sequence {r}")
    finally:
        del x, y
        torch.cuda.empty_cache()


def test_real_past_2_31_elements(gpu):
    """R2C then C2R at N = 6000, batch 357914: 2^31 + 352 reals on the real side,
    rows sampled at the marks of both the real and the half-complex buffers."""
    import torch

    n, batch = 6000, 357914
    h = n // 2 + 1
    assert n * batch > 1 << 31
    gen = torch.Generator(device=gpu)
    gen.manual_seed(n + 1)
    x = torch.rand(n * batch, dtype=torch.float32, device=gpu, generator=gen) * 2 - 1
    X = torch.empty(h * batch, dtype=torch.complex64, device=gpu)
    z = torch.empty_like(x)
    try:
        with ff.Plan.one_d(n, ff.Type.R2C, batch) as pf, ff.Plan.one_d(n, ff.Type.C2R, batch) as pi:
            pf.set_stream(stream_of(gpu))
            pi.set_stream(stream_of(gpu))
            pf.r2c(x.data_ptr(), X.data_ptr())
            pi.c2r(X.data_ptr(), z.data_ptr())
            sync(gpu)
        for r in sorted(set(_rows(n, batch, n)) | set(_rows(n, batch, h))):
            xr = host(x[r * n:(r + 1) * n]).astype(np.float64)
            check_accuracy(host(X[r * h:(r + 1) * h])[None], ora.rfft(xr[None]), ff.Type.R2C, n, row_len=h,
                           what=f"R2C sequence {r}")
            check_accuracy(host(z[r * n:(r + 1) * n])[None], xr[None] * n, ff.Type.C2R, n, row_len=n, factor=2.0,
                           what=f"roundtrip sequence {r}")
    finally:
        del x, X, z
        torch.cuda.empty_cache()


def test_2d_batch_past_2_31_elements(gpu):
    """Rank-2 C2C 2048 x 2048, batch 513 (flagfftPlanMany): rows then the
    column pass over column-interleaved units, batch entries sampled either
    side of the 2^29- and 2^31-element marks."""
    import torch

    n0 = n1 = 2048
    batch = 513
    m = n0 * n1
    assert m * batch > 1 << 31
    gen = torch.Generator(device=gpu)
    gen.manual_seed(m + batch)
    x = torch.randn(m * batch, dtype=torch.complex64, device=gpu, generator=gen)
    y = torch.empty_like(x)
    try:
        with ff.Plan.many(2, [n0, n1], None, 1, m, None, 1, m, ff.Type.C2C, batch) as p:
            p.set_stream(stream_of(gpu))
            p.c2c(x.data_ptr(), y.data_ptr(), ff.FORWARD)
            sync(gpu)
        for r in _rows(m, batch, m):
            xr = host(x[r * m:(r + 1) * m]).astype(np.complex128).reshape(1, n0, n1)
            check_accuracy(host(y[r * m:(r + 1) * m]).reshape(1, n0, n1), ora.fft2(xr), ff.Type.C2C, m, row_len=m,
                           what=f"2D entry {r}")
    finally:
        del x, y
        torch.cuda.empty_cache()


def _closed_form_err(y, n, fn, chunk=1 << 26):
    """Relative L2 of y against fn(k) over k < n, FP64 on the device in chunks."""
    import torch

    num = den = 0.0
    for s in range(0, n, chunk):
        k = torch.arange(s, min(n, s + chunk), dtype=torch.int64, device=y.device)
        ref = fn(k, s)
        d = y[s:s + k.numel()].to(torch.complex128) - ref
        num += float((d.real * d.real + d.imag * d.imag).sum())
        den += float((ref.real * ref.real + ref.imag * ref.imag).sum())
    return (num / den) ** 0.5


@pytest.mark.parametrize("logn", [28, 30])
def test_longest_chains(gpu, logn):
    """N = 2^28 and 2^30 (three-pass chains, one sequence of 2 / 8 GiB): the
    spectrum of an impulse at j is w_N^(jk) (every inter-pass twiddle enters
    it), and forward then inverse of N(0,1) data is N x.  The expected values
    are closed forms in FP64 on the device; the FP64 oracle would need minutes
    at this length."""
    import math

    import torch

    n = 1 << logn
    bound = 2e-6 * logn
    j = 123456789 % n
    x = torch.zeros(n, dtype=torch.complex64, device=gpu)
    x[j] = 1
    y = torch.empty_like(x)
    try:
        with ff.Plan.one_d(n, ff.Type.C2C, 1) as p:
            assert "passes=3" in p.description()
            p.set_stream(stream_of(gpu))
            p.c2c(x.data_ptr(), y.data_ptr(), ff.FORWARD)
            sync(gpu)

            def impulse(k, s):
                ang = ((j * k) % n).to(torch.float64) * (-2.0 * math.pi / n)
                return torch.complex(torch.cos(ang), torch.sin(ang))

            e = _closed_form_err(y, n, impulse)
            assert e <= bound, f"impulse at {j}: rel L2 {e:.3e} > {bound:.3e}"
            gen = torch.Generator(device=gpu)
            gen.manual_seed(logn)
            x = torch.randn(n, dtype=torch.complex64, device=gpu, generator=gen)
            p.c2c(x.data_ptr(), y.data_ptr(), ff.FORWARD)
            p.c2c(y.data_ptr(), y.data_ptr(), ff.INVERSE)
            sync(gpu)
            e = _closed_form_err(y, n, lambda k, s: x[s:s + k.numel()].to(torch.complex128) * n)
            assert e <= 2 * bound, f"roundtrip: rel L2 {e:.3e} > {2 * bound:.3e}"
    finally:
        del x, y
        torch.cuda.empty_cache()
