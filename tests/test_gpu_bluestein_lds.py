"""LDS-resident Bluestein kernels against the oracle (GPU).

Covers every instance the planner can pick for a prime length whose
convolution does not fit the generic single-pass kernel:
  FP32, whole convolution in LDS     M = 16384 (n <= 8192), 20160 (n <= 10080)
  FP32, two halves of M / 2 in LDS   M = 32768 (n <= 16384), 40320 (n <= 20160)
  FP64, whole convolution in LDS     M = 8192 (n <= 4096), 10080 (n <= 5040)
  FP64, two halves                   M = 16384 (n <= 8192), 20160 (n <= 10080)
each forward and inverse, batch 1 and 3, out-of-place and in-place, the FP32
kernel variants chosen through the units tuning knob, strided sequences (2D
column passes) and real / Hermitian sources (odd R2C / C2R lengths).
Bounds: FP32 2e-6 log2 N, FP64 1e-14 log2 N relative L2 per transform.
"""
import numpy as np
import pytest

from flagfft_amd import flagfft as ff
from oracle import oracle as ora

from .helpers import check_accuracy, complex_input, dev_tensor, host, real_input, run_c2c, stream_of, sync

pytestmark = pytest.mark.gpu

# (n, expected description fragment)
F32 = [(4099, "conv_length=16384, fused=1, lds_resident=1)"), (9001, "conv_length=20160, fused=1, lds_resident=1)"),
       (12011, "conv_length=32768, fused=1, lds_resident=1, halves=1"),
       (20011, "conv_length=40320, fused=1, lds_resident=1, halves=1")]
F64 = [(2063, "conv_length=8192, fused=1, lds_resident=1)"), (4099, "conv_length=10080, fused=1, lds_resident=1)"),
       (7919, "conv_length=16384, fused=1, lds_resident=1, halves=1"),
       (10007, "conv_length=20160, fused=1, lds_resident=1, halves=1")]


def _describe(n, t, batch):
    r, text = ff.describe_host(1, [n], t, batch)
    ff.check("flagfftxPlanDescribe", r)
    return text


def _check(dev, n, dtype, batch, direction, inplace):
    t = ff.Type.Z2Z if dtype == np.complex128 else ff.Type.C2C
    x = complex_input(ora.accuracy_seed(int(t), n, batch, 3), (batch, n), dtype, "normal")
    y = run_c2c(dev, x, direction, n, batch, inplace=inplace)
    check_accuracy(y, ora.fft(x.astype(np.complex128), sign=direction), t, n, row_len=n,
                   what=f"dir={direction} batch={batch} inplace={inplace}")


@pytest.mark.parametrize("n,frag", F32, ids=[str(n) for n, _ in F32])
def test_fp32_instances(gpu, n, frag):
    assert frag in _describe(n, ff.Type.C2C, 3)
    for batch in (1, 3):
        for direction in (ff.FORWARD, ff.INVERSE):
            _check(gpu, n, np.complex64, batch, direction, inplace=False)
    _check(gpu, n, np.complex64, 3, ff.FORWARD, inplace=True)


@pytest.mark.parametrize("n,frag", F64, ids=[str(n) for n, _ in F64])
def test_fp64_instances(gpu, n, frag):
    assert frag in _describe(n, ff.Type.Z2Z, 3)
    for batch in (1, 3):
        for direction in (ff.FORWARD, ff.INVERSE):
            _check(gpu, n, np.complex128, batch, direction, inplace=False)
    _check(gpu, n, np.complex128, 3, ff.INVERSE, inplace=True)


@pytest.mark.parametrize("n", [4099, 10007])
def test_fp32_kernel_variants(gpu, n):
    """units 1..4: Stockham (prefetched / per-stage twiddles), in-place (table / binary-powered twiddles)."""
    x = complex_input(ora.accuracy_seed(int(ff.Type.C2C), n, 2, 5), (2, n), np.complex64, "normal")
    ref = ora.fft(x.astype(np.complex128), sign=ff.FORWARD)
    outs = []
    try:
        for units in (1, 2, 3, 4):
            ff.set_tuning(units=units)
            y = run_c2c(gpu, x, ff.FORWARD, n, 2)
            check_accuracy(y, ref, ff.Type.C2C, n, row_len=n, what=f"units={units}")
            outs.append(y)
    finally:
        ff.set_tuning()
    # the four kernels compute the same convolution: they agree to FP32 rounding
    for y in outs[1:]:
        assert ora.rel_l2(y, outs[0], n)[0] < 2e-6


@pytest.mark.parametrize("shape,dbl", [((4099, 4), False), ((12011, 2), False), ((2053, 3), True), ((7919, 2), True)],
                         ids=["4099x4", "12011x2", "z2z-2053x3", "z2z-7919x2"])
def test_strided_columns(gpu, shape, dbl):
    """2D transforms whose column pass is an LDS Bluestein over strided sequences."""
    n0, n1 = shape
    t = ff.Type.Z2Z if dbl else ff.Type.C2C
    x = complex_input(n0 + 13 * n1, (1, n0, n1), np.complex128 if dbl else np.complex64, "normal")
    xd = dev_tensor(x, gpu)
    yd = xd.clone().zero_()
    with ff.Plan.two_d(n0, n1, t) as p:
        p.set_stream(stream_of(gpu))
        (p.z2z if dbl else p.c2c)(xd.data_ptr(), yd.data_ptr(), ff.FORWARD)
        sync(gpu)
    n = n0 * n1
    check_accuracy(host(yd), ora.fft2(x.astype(np.complex128)), t, n, row_len=n, factor=2.0, what="2D column pass")


@pytest.mark.parametrize("n,dbl", [(9001, False), (12011, False), (4099, True), (7919, True)],
                         ids=["r2c-9001", "r2c-12011", "d2z-4099", "d2z-7919"])
def test_real_sources(gpu, n, dbl):
    """Odd real lengths run a complex Bluestein core with real / Hermitian sources."""
    tf, ti = (ff.Type.D2Z, ff.Type.Z2D) if dbl else (ff.Type.R2C, ff.Type.C2R)
    batch = 2
    x = real_input(ora.accuracy_seed(int(tf), n, batch, 7), (batch, n), np.float64 if dbl else np.float32)
    xd = dev_tensor(x, gpu)
    Xd = dev_tensor(np.zeros((batch, n // 2 + 1), np.complex128 if dbl else np.complex64), gpu)
    zd = xd.clone().zero_()
    with ff.Plan.one_d(n, tf, batch) as pf, ff.Plan.one_d(n, ti, batch) as pi:
        pf.set_stream(stream_of(gpu))
        pi.set_stream(stream_of(gpu))
        (pf.d2z if dbl else pf.r2c)(xd.data_ptr(), Xd.data_ptr())
        (pi.z2d if dbl else pi.c2r)(Xd.data_ptr(), zd.data_ptr())
        sync(gpu)
    check_accuracy(host(Xd), ora.rfft(x.astype(np.float64)), tf, n, row_len=n // 2 + 1, what="R2C")
    check_accuracy(host(zd), x.astype(np.float64) * n, ti, n, row_len=n, factor=2.0, what="roundtrip")


SPLITS = [(4099, False), (9001, False), (12011, False), (20011, False),
          (2053, True), (4099, True), (7919, True), (10007, True)]


@pytest.mark.parametrize("n,dbl", SPLITS, ids=[f"{'z2z' if d else 'c2c'}-{n}" for n, d in SPLITS])
def test_split_forms(gpu, n, dbl):
    """Convolution spread over 2 / 4 workgroups per sequence (units 6 / 7) plus
    the combine launch: both directions, several execs in a row, in place and
    as graph replays (bit-identical to the first exec)."""
    t = ff.Type.Z2Z if dbl else ff.Type.C2C
    dtype = np.complex128 if dbl else np.complex64
    for batch in (3, 8):  # 8: the XCD-aligned block order
        x = complex_input(ora.accuracy_seed(int(t), n, batch, 11), (batch, n), dtype, "normal")
        ref_f = ora.fft(x.astype(np.complex128), sign=ff.FORWARD)
        ref_i = ora.fft(x.astype(np.complex128), sign=ff.INVERSE)
        for units, p in ((6, 2), (7, 4)):
            ff.set_tuning(units=units)
            try:
                plan = ff.Plan.one_d(n, t, batch)
            finally:
                ff.set_tuning()
            with plan:
                desc = plan.description()
                if f"k_bluestein_split<{'fp64' if dbl else 'fp32'}" not in desc:
                    continue  # no instance of this split for this M
                assert f"k_bluestein_combine<{'fp64' if dbl else 'fp32'},{p}>" in desc
                plan.set_stream(stream_of(gpu))
                ex = plan.z2z if dbl else plan.c2c
                xd = dev_tensor(x, gpu)
                outs = []
                for rep in range(3):
                    yd = xd.clone().zero_()
                    ex(xd.data_ptr(), yd.data_ptr(), ff.FORWARD if rep != 1 else ff.INVERSE)
                    outs.append(yd)
                sync(gpu)
                for rep, yd in enumerate(outs):
                    check_accuracy(host(yd), ref_i if rep == 1 else ref_f, t, n, row_len=n,
                                   what=f"units={units} batch={batch} exec {rep}")
                assert np.array_equal(host(outs[0]), host(outs[2]))
                zd = xd.clone()
                ex(zd.data_ptr(), zd.data_ptr(), ff.FORWARD)
                sync(gpu)
                assert np.array_equal(host(zd), host(outs[0])), "in place"
                plan.set_graph_mode(True)
                for rep in range(3):
                    gd = xd.clone().zero_()
                    ex(xd.data_ptr(), gd.data_ptr(), ff.FORWARD)
                    sync(gpu)
                    assert np.array_equal(host(gd), host(outs[0])), f"graph replay {rep}"
