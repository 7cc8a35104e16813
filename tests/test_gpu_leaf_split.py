"""Single-pass transforms of few sequences spread over P workgroups (GPU).

k_leaf_split computes X[P k + h] = FFT_L(y_h)[k] with
y_h[n1] = w_N^(n1 h) sum_n2 x[n1 + L n2] w_P^(n2 h) (N = P L) on workgroup
(sequence, h).  The planner takes it for contiguous complex batches smaller
than half the CU count when the length has an instance (8192, 16384 and
10935 = 5 x 3^7), in place of one workgroup per sequence or a multi-pass
chain.  Checked against the FP64 oracle within the stated bounds, both
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
