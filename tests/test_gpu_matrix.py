"""The reference's accuracy matrix (conf/test_matrix.yaml, expanded as
tools/run_tests.py:156-206 does) against the FP64 oracle (GPU).

`full`: every Cooley-Tukey size and every Bluestein/Rader size, batches
{1, 4, 256}, input scales {2^-24, 1, 2^24}, for each 1D operator of
conf/operators.yaml (C2C / Z2Z forward and inverse, R2C, D2Z, C2R, Z2D and the
R2C->C2R / D2Z->Z2D roundtrips).  `2d_full`: the 2D sizes with the same
batches and scales for C2C / Z2Z (both directions), R2C / D2Z and C2R / Z2D.
The values of the matrix are restated here (conf/test_matrix.yaml:8-93).

Inputs are iid Uniform[-1, 1) drawn on the device from a seeded generator;
the oracle checks every transform of batches 1 and 4 and a sample of four
transforms (first, second, middle, last) of batch 256.  Scales are powers of
two, so the oracle of the scaled input is the scaled oracle of the input.
Bounds: the reference's accuracy model and FP32 2e-6 log2 N / FP64
1e-14 log2 N relative L2 per transform; roundtrips N x within twice that.
"""
import numpy as np
import pytest

from flagfft_amd import flagfft as ff
from oracle import oracle as ora

from .helpers import check_accuracy, stream_of, sync
from .test_gpu_2d import _plan as _plan2d

pytestmark = pytest.mark.gpu

# conf/test_matrix.yaml:8-40
SIZES_CT = [16, 64, 81, 243, 256, 361, 512, 2048, 4096, 8192, 16384]
SIZES_BS = [23, 67, 997, 1009]
BATCHES = [1, 4, 256]
SCALES = [2.0 ** -24, 1.0, 2.0 ** 24]
# conf/test_matrix.yaml:81-93 (2d_full)
SIZES_2D = [(16, 16), (64, 64), (128, 256), (256, 128), (256, 256), (512, 512), (23, 23), (997, 16), (16, 997)]

OPS_1D = ["c2c", "z2z", "r2c", "d2z", "c2r", "z2d", "r2c_c2r", "d2z_z2d"]


def _rows(batch: int):
    return list(range(batch)) if batch <= 4 else [0, 1, batch // 2, batch - 1]


def _uniform(dev, shape, dbl: bool, cplx: bool, seed: int):
    """iid Uniform[-1, 1) on the device (real and imaginary parts independent)."""
    import torch

    g = torch.Generator(device=dev)
    g.manual_seed(seed)
    rdt = torch.float64 if dbl else torch.float32
    if cplx:
        v = torch.rand(tuple(shape) + (2,), dtype=rdt, device=dev, generator=g) * 2 - 1
        return torch.view_as_complex(v.contiguous())
    return torch.rand(tuple(shape), dtype=rdt, device=dev, generator=g) * 2 - 1


def _sample(t, rows):
    return t[rows].cpu().numpy()


@pytest.mark.parametrize("n", SIZES_CT + SIZES_BS)
@pytest.mark.parametrize("op", OPS_1D)
def test_full_1d(gpu, op, n):
    import torch

    dbl = op in ("z2z", "d2z", "z2d", "d2z_z2d")
    cdt = torch.complex128 if dbl else torch.complex64
    rdt = torch.float64 if dbl else torch.float32
    h = n // 2 + 1
    seed = 0x4D41 + 131 * n + OPS_1D.index(op)
    if op in ("c2c", "z2z"):
        t = ff.Type.Z2Z if dbl else ff.Type.C2C
        x = _uniform(gpu, (256, n), dbl, True, seed)
        xs = x.cpu().numpy().astype(np.complex128)
        refs = {d: ora.fft(xs, sign=d) for d in (ff.FORWARD, ff.INVERSE)}
        for batch in BATCHES:
            rows = _rows(batch)
            with ff.Plan.one_d(n, t, batch) as p:
                p.set_stream(stream_of(gpu))
                ex = p.z2z if dbl else p.c2c
                for scale in SCALES:
                    xi = (x[:batch] * scale).contiguous()
                    y = torch.empty_like(xi)
                    for d in (ff.FORWARD, ff.INVERSE):
                        ex(xi.data_ptr(), y.data_ptr(), d)
                        sync(gpu)
                        check_accuracy(_sample(y, rows), refs[d][rows] * scale, t, n, row_len=n,
                                       what=f"{op} dir={d} batch={batch} scale={scale}")
    elif op in ("r2c", "d2z"):
        t = ff.Type.D2Z if dbl else ff.Type.R2C
        x = _uniform(gpu, (256, n), dbl, False, seed)
        ref = ora.rfft(x.cpu().numpy().astype(np.float64))
        for batch in BATCHES:
            rows = _rows(batch)
            with ff.Plan.one_d(n, t, batch) as p:
                p.set_stream(stream_of(gpu))
                for scale in SCALES:
                    xi = (x[:batch] * scale).contiguous()
                    y = torch.empty((batch, h), dtype=cdt, device=gpu)
                    (p.d2z if dbl else p.r2c)(xi.data_ptr(), y.data_ptr())
                    sync(gpu)
                    check_accuracy(_sample(y, rows), ref[rows] * scale, t, n, row_len=h,
                                   what=f"{op} batch={batch} scale={scale}")
    elif op in ("c2r", "z2d"):
        t = ff.Type.Z2D if dbl else ff.Type.C2R
        X = _uniform(gpu, (256, h), dbl, True, seed)
        ref = ora.irfft(X.cpu().numpy().astype(np.complex128), n)
        for batch in BATCHES:
            rows = _rows(batch)
            with ff.Plan.one_d(n, t, batch) as p:
                p.set_stream(stream_of(gpu))
                for scale in SCALES:
                    Xi = (X[:batch] * scale).contiguous()
                    keep = Xi.clone()
                    y = torch.empty((batch, n), dtype=rdt, device=gpu)
                    (p.z2d if dbl else p.c2r)(Xi.data_ptr(), y.data_ptr())
                    sync(gpu)
                    assert torch.equal(Xi, keep), "C2R must not modify its input"
                    check_accuracy(_sample(y, rows), ref[rows] * scale, t, n, row_len=n,
                                   what=f"{op} batch={batch} scale={scale}")
    else:
        tf, ti = (ff.Type.D2Z, ff.Type.Z2D) if dbl else (ff.Type.R2C, ff.Type.C2R)
        x = _uniform(gpu, (256, n), dbl, False, seed)
        xs = x.cpu().numpy().astype(np.float64)
        for batch in BATCHES:
            rows = _rows(batch)
            with ff.Plan.one_d(n, tf, batch) as pf, ff.Plan.one_d(n, ti, batch) as pi:
                pf.set_stream(stream_of(gpu))
                pi.set_stream(stream_of(gpu))
                for scale in SCALES:
                    xi = (x[:batch] * scale).contiguous()
                    X = torch.empty((batch, h), dtype=cdt, device=gpu)
                    z = torch.empty_like(xi)
                    (pf.d2z if dbl else pf.r2c)(xi.data_ptr(), X.data_ptr())
                    (pi.z2d if dbl else pi.c2r)(X.data_ptr(), z.data_ptr())
                    sync(gpu)
                    check_accuracy(_sample(z, rows), xs[rows] * (scale * n), ti, n, row_len=n, factor=2.0,
                                   what=f"{op} roundtrip batch={batch} scale={scale}")


@pytest.mark.parametrize("dbl", [False, True], ids=["fp32", "fp64"])
@pytest.mark.parametrize("shape", SIZES_2D, ids=lambda s: f"{s[0]}x{s[1]}")
def test_full_2d(gpu, shape, dbl):
    """2d_full: C2C / Z2Z forward and inverse, R2C / D2Z, C2R / Z2D."""
    import torch

    n0, n1 = shape
    n, h = n0 * n1, n1 // 2 + 1
    tc = ff.Type.Z2Z if dbl else ff.Type.C2C
    tf = ff.Type.D2Z if dbl else ff.Type.R2C
    ti = ff.Type.Z2D if dbl else ff.Type.C2R
    cdt = torch.complex128 if dbl else torch.complex64
    rdt = torch.float64 if dbl else torch.float32
    seed = 0x2D00 + 7 * n0 + n1 + (1 << 20) * int(dbl)
    xc = _uniform(gpu, (256, n0, n1), dbl, True, seed)
    xr = _uniform(gpu, (256, n0, n1), dbl, False, seed + 1)
    Xh = _uniform(gpu, (256, n0, h), dbl, True, seed + 2)
    for batch in BATCHES:
        rows = _rows(batch)
        xcs = xc[rows].cpu().numpy().astype(np.complex128)
        refc = {ff.FORWARD: ora.fft2(xcs, sign=ff.FORWARD), ff.INVERSE: ora.fft2(xcs, sign=ff.INVERSE)}
        refr = ora.rfft2(xr[rows].cpu().numpy().astype(np.float64))
        refh = ora.irfft2(Xh[rows].cpu().numpy().astype(np.complex128), n1)
        with _plan2d(n0, n1, tc, batch) as pc, _plan2d(n0, n1, tf, batch) as pf, _plan2d(n0, n1, ti, batch) as pi:
            for p in (pc, pf, pi):
                p.set_stream(stream_of(gpu))
            for scale in SCALES:
                what = f"2D {n0}x{n1} batch={batch} scale={scale}"
                xi = (xc[:batch] * scale).contiguous()
                y = torch.empty_like(xi)
                for d in (ff.FORWARD, ff.INVERSE):
                    (pc.z2z if dbl else pc.c2c)(xi.data_ptr(), y.data_ptr(), d)
                    sync(gpu)
                    check_accuracy(_sample(y, rows).reshape(len(rows), n), refc[d].reshape(len(rows), n) * scale,
                                   tc, n, row_len=n, factor=2.0, what=f"{what} C2C dir={d}")
                ri = (xr[:batch] * scale).contiguous()
                Y = torch.empty((batch, n0, h), dtype=cdt, device=gpu)
                (pf.d2z if dbl else pf.r2c)(ri.data_ptr(), Y.data_ptr())
                sync(gpu)
                check_accuracy(_sample(Y, rows).reshape(len(rows), -1), refr.reshape(len(rows), -1) * scale, tf, n,
                               row_len=n0 * h, factor=2.0, what=f"{what} R2C")
                hi = (Xh[:batch] * scale).contiguous()
                z = torch.empty((batch, n0, n1), dtype=rdt, device=gpu)
                (pi.z2d if dbl else pi.c2r)(hi.data_ptr(), z.data_ptr())
                sync(gpu)
                check_accuracy(_sample(z, rows).reshape(len(rows), n), refh.reshape(len(rows), n) * scale, ti, n,
                               row_len=n, factor=2.0, what=f"{what} C2R")
