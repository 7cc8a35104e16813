"""Single-pass transforms of few sequences spread over P workgroups (GPU).

k_leaf_split computes X[P k + h] = FFT_L(y_h)[k] with
y_h[n1] = w_N^(n1 h) sum_n2 x[n1 + L n2] w_P^(n2 h) (N = P L) on workgroup
(sequence, h).  The planner takes it for contiguous complex batches whose
split workgroups (batch x P) are at most twice the CU count when the length
has an instance (8192, 16384 and 10935 = 5 x 3^7), in place of one
workgroup per sequence or a multi-pass chain.  Checked against the FP64 oracle within the stated bounds, both
directions, in place, with a sequence count that is and is not a multiple of
8 (the XCD-aligned and the plain workgroup order), and against the non-split
plan of the same length.
"""
import numpy as np
import pytest

from flagfft_amd import flagfft as ff
from oracle import oracle as ora

from .helpers import check_accuracy, complex_input, run_c2c

pytestmark = pytest.mark.gpu

CASES = [(10935, 64), (10935, 3), (8192, 8), (8192, 5), (16384, 1), (16384, 16)]


@pytest.fixture(autouse=True)
def untuned(monkeypatch):
    """The planner's own choice (tuned-plan rows may prefer another form)."""
    monkeypatch.setenv("FLAGFFT_TUNE_DISABLE", "1")


def _uses_split(n, t, batch):
    with ff.Plan.one_d(n, t, batch) as p:
        return "k_leaf_split<" in p.description()


@pytest.mark.parametrize("dtype", [np.complex64, np.complex128], ids=["c2c", "z2z"])
@pytest.mark.parametrize("n,batch", CASES, ids=[f"{n}x{b}" for n, b in CASES])
def test_split_leaf(gpu, n, batch, dtype):
    t = ff.Type.Z2Z if dtype == np.complex128 else ff.Type.C2C
    assert _uses_split(n, t, batch)
    with ff.Plan.one_d(n, t, batch) as p:
        # in-place calls run an executable without the split (each of the
        # split's workgroups reads the whole sequence)
        ip = p.description().split("-- Forward execution, in place --")
        assert len(ip) == 2 and "k_leaf_split<" not in ip[1].split("-- Inverse")[0]
    x = complex_input(ora.accuracy_seed(int(t), n, batch, 9), (batch, n), dtype, "normal")
    for direction, inplace in ((ff.FORWARD, False), (ff.INVERSE, False), (ff.FORWARD, True)):
        y = run_c2c(gpu, x, direction, n, batch, inplace=inplace)
        check_accuracy(y, ora.fft(x.astype(np.complex128), sign=direction), t, n, row_len=n,
                       what=f"dir={direction} batch={batch} inplace={inplace}")


def test_split_matches_unsplit(gpu):
    """The split and the one-workgroup-per-sequence leaf agree to FP32 rounding."""
    n, batch = 10935, 13
    x = complex_input(77, (batch, n), np.complex64, "normal")
    split = run_c2c(gpu, x, ff.FORWARD, n, batch)
    try:
        ff.set_tuning(units=1)
        whole = run_c2c(gpu, x, ff.FORWARD, n, batch)
    finally:
        ff.set_tuning()
    assert ora.rel_l2(split, whole, n)[0] < 2e-6


REAL_CASES = [(6000, 128), (6000, 5), (6000, 1)]


@pytest.mark.parametrize("dbl", [False, True], ids=["fp32", "fp64"])
@pytest.mark.parametrize("n,batch", REAL_CASES, ids=[f"{n}x{b}" for n, b in REAL_CASES])
def test_split_leaf_real(gpu, n, batch, dbl):
    """Real rows of 6000 (3000-point core split in two, pack / unpack fused): R2C
    against the oracle's half spectrum, C2R of a non-Hermitian input against the
    oracle (Im of DC / Nyquist ignored), R2C -> C2R = n x."""
    import torch

    from .helpers import dev_tensor, host, real_input, stream_of, sync

    tf, ti = (ff.Type.D2Z, ff.Type.Z2D) if dbl else (ff.Type.R2C, ff.Type.C2R)
    assert _uses_split(n, tf, batch) and _uses_split(n, ti, batch)
    h = n // 2 + 1
    cdt = np.complex128 if dbl else np.complex64
    x = real_input(ora.accuracy_seed(int(tf), n, batch, 11), (batch, n), np.float64 if dbl else np.float32)
    X = complex_input(ora.accuracy_seed(int(ti), n, batch, 12), (batch, h), cdt)
    xd = dev_tensor(x, gpu)
    Xd = torch.zeros((batch, h), dtype=torch.complex128 if dbl else torch.complex64, device=gpu)
    Xin = dev_tensor(X, gpu)
    yd = xd.clone().zero_()
    zd = xd.clone().zero_()
    with ff.Plan.one_d(n, tf, batch) as pf, ff.Plan.one_d(n, ti, batch) as pi:
        pf.set_stream(stream_of(gpu))
        pi.set_stream(stream_of(gpu))
        (pf.d2z if dbl else pf.r2c)(xd.data_ptr(), Xd.data_ptr())
        (pi.z2d if dbl else pi.c2r)(Xd.data_ptr(), zd.data_ptr())
        (pi.z2d if dbl else pi.c2r)(Xin.data_ptr(), yd.data_ptr())
        sync(gpu)
    check_accuracy(host(Xd), ora.rfft(x.astype(np.float64)), tf, n, row_len=h, what=f"R2C batch={batch}")
    check_accuracy(host(yd), ora.irfft(X.astype(np.complex128), n), ti, n, row_len=n, what=f"C2R batch={batch}")
    check_accuracy(host(zd), x.astype(np.float64) * n, ti, n, row_len=n, factor=2.0, what="roundtrip")
