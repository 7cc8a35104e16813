"""Split-precision matrix-core C2C (N=1024) against the FP64 oracle (GPU).

The radix-32 x 32 stages run as BF16 / FP16 hi+lo matrix products with FP32
accumulation (flagfftxPlan1dCompute).  Each test states the error it measures
against the FP64 oracle next to the plain FP32 vector path on the same data.

Bounds (relative L2 against the FP64 oracle, N=1024):
  FP16 hi+lo: 2e-6 * log2 N = 2e-5, the stated FP32 bound (the split keeps 22
              significand bits, the dropped lo*lo term is ~2^-22 relative);
  BF16 hi+lo: 4e-5 (16 significand bits, dropped term ~2^-16 relative).
"""
import math

import numpy as np
import pytest

from flagfft_amd import flagfft as ff
from oracle import oracle as ora

from .helpers import complex_input, dev_tensor, host, stream_of, sync

pytestmark = pytest.mark.gpu

N = 1024
BOUND = {ff.COMPUTE_FP16X3: 2e-6 * math.log2(N), ff.COMPUTE_BF16X3: 4e-5}
NAMES = {ff.COMPUTE_FP16X3: "fp16x3", ff.COMPUTE_BF16X3: "bf16x3"}


def _run(dev, x, direction, compute):
    batch = x.shape[0]
    xd = dev_tensor(x, dev)
    yd = xd.clone().zero_()
    with ff.Plan.one_d_compute(N, ff.Type.C2C, batch, compute) as p:
        p.set_stream(stream_of(dev))
        p.c2c(xd.data_ptr(), yd.data_ptr(), direction)
        sync(dev)
        desc = p.description()
    return host(yd), desc


@pytest.mark.parametrize("compute", [ff.COMPUTE_FP16X3, ff.COMPUTE_BF16X3], ids=["fp16x3", "bf16x3"])
@pytest.mark.parametrize("batch", [1, 7, 256])
@pytest.mark.parametrize("direction", [ff.FORWARD, ff.INVERSE], ids=["fwd", "inv"])
def test_split_precision_matches_oracle(gpu, compute, batch, direction):
    seed = ora.accuracy_seed(int(ff.Type.C2C), N, batch, 7)
    x = complex_input(seed, (batch, N), np.complex64, "normal")
    y, desc = _run(gpu, x, direction, compute)
    assert "MatrixCoreSplit(n=1024" in desc and "k_mfma1024" in desc
    ref = ora.fft(x.astype(np.complex128), sign=direction)
    l2, linf = ora.rel_l2(y, ref, N)
    print(f"{NAMES[compute]} batch={batch} dir={direction}: rel_l2={l2:.3e} rel_linf={linf:.3e} "
          f"(bound {BOUND[compute]:.1e})")
    assert np.isfinite(l2) and l2 <= BOUND[compute]


@pytest.mark.parametrize("compute", [ff.COMPUTE_FP16X3, ff.COMPUTE_BF16X3], ids=["fp16x3", "bf16x3"])
@pytest.mark.parametrize("scale", [2.0 ** -20, 2.0 ** 20, 1e30])
def test_split_precision_input_scales(gpu, compute, scale):
    """Per-transform power-of-two range scaling keeps FP16 usable at any magnitude."""
    x = complex_input(ora.accuracy_seed(int(ff.Type.C2C), N, 4, 8), (4, N), np.complex64, "normal", scale)
    y, _ = _run(gpu, x, ff.FORWARD, compute)
    ref = ora.fft(x.astype(np.complex128))
    l2, _ = ora.rel_l2(y, ref, N)
    assert np.isfinite(l2) and l2 <= BOUND[compute], f"scale={scale} rel_l2={l2:.3e}"


@pytest.mark.parametrize("compute", [ff.COMPUTE_FP16X3, ff.COMPUTE_BF16X3], ids=["fp16x3", "bf16x3"])
def test_split_precision_closed_forms(gpu, compute):
    """Impulse at j -> w^(jk); constant -> N delta_0; a pure tone -> N delta_m."""
    rows = []
    for j in (0, 1, 33, 1023):
        e = np.zeros(N, np.complex64)
        e[j] = 1
        rows.append(e)
    rows.append(np.ones(N, np.complex64))
    rows.append(np.exp(2j * np.pi * 77 * np.arange(N) / N).astype(np.complex64))
    x = np.stack(rows)
    y, _ = _run(gpu, x, ff.FORWARD, compute)
    k = np.arange(N)
    expect = [np.exp(-2j * np.pi * j * k / N) for j in (0, 1, 33, 1023)]
    expect.append(N * (k == 0))
    expect.append(N * (k == 77))
    for got, want in zip(y, expect):
        assert np.linalg.norm(got - want) / np.linalg.norm(want) <= BOUND[compute]


def test_split_precision_same_data_as_vector_path(gpu):
    """Error of each path on the same data (the report the precision form asks for)."""
    batch = 512
    x = complex_input(ora.accuracy_seed(int(ff.Type.C2C), N, batch, 9), (batch, N), np.complex64, "normal")
    ref = ora.fft(x.astype(np.complex128))
    errs = {}
    for compute in (ff.COMPUTE_VECTOR, ff.COMPUTE_FP16X3, ff.COMPUTE_BF16X3):
        y, _ = _run(gpu, x, ff.FORWARD, compute)
        errs[compute] = ora.rel_l2(y, ref, N)[0]
    print("rel_l2 vs FP64 oracle: " + ", ".join(f"{k}={v:.3e}" for k, v in errs.items()))
    assert errs[ff.COMPUTE_VECTOR] <= 2e-6 * math.log2(N)
    assert errs[ff.COMPUTE_FP16X3] <= BOUND[ff.COMPUTE_FP16X3]
    assert errs[ff.COMPUTE_BF16X3] <= BOUND[ff.COMPUTE_BF16X3]


# Rows sampled from every part of a 65536-transform batch: the first wave
# iteration (rows below 4 x 1024 groups), the second and later iterations of
# each wave's loop (rows past 4096, read by the software-pipelined prefetch),
# and the last transforms of the batch.
FULL_BATCH = 65536
SAMPLE_ROWS = np.r_[0:64, 4096:4160, 8192:8200, 32768:32832, 65472:65536]


def _device_normal(dev, batch, seed):
    """Complex N(0,1) rows (re and im iid N(0,1)) generated on the device."""
    import torch

    g = torch.Generator(device=dev)
    g.manual_seed(seed)
    return torch.randn((batch, N), dtype=torch.complex64, device=dev, generator=g) * (2.0 ** 0.5)


@pytest.mark.parametrize("direction", [ff.FORWARD, ff.INVERSE], ids=["fwd", "inv"])
def test_full_configuration_1024x65536_all_paths(gpu, direction):
    """N=1024 x 65536 (the precision-form configuration) on the same data for the
    FP32 vector path, FP16 hi+lo and BF16 hi+lo: sampled rows against the FP64
    oracle, every row of each split path against the vector path."""
    import torch

    xd = _device_normal(gpu, FULL_BATCH, 1024 + direction)
    x = xd[SAMPLE_ROWS].cpu().numpy().astype(np.complex128)
    ref = ora.fft(x, sign=direction)
    outs, errs = {}, {}
    for compute in (ff.COMPUTE_VECTOR, ff.COMPUTE_FP16X3, ff.COMPUTE_BF16X3):
        yd = torch.empty_like(xd)
        with ff.Plan.one_d_compute(N, ff.Type.C2C, FULL_BATCH, compute) as p:
            p.set_stream(stream_of(gpu))
            p.c2c(xd.data_ptr(), yd.data_ptr(), direction)
            sync(gpu)
            if compute != ff.COMPUTE_VECTOR:
                assert "k_mfma1024" in p.description()
        errs[compute] = ora.rel_l2(yd[SAMPLE_ROWS].cpu().numpy(), ref, N)[0]
        outs[compute] = yd
    bound = {ff.COMPUTE_VECTOR: 2e-6 * math.log2(N), **BOUND}
    print(f"N=1024 x {FULL_BATCH} dir={direction} rel_l2 vs FP64 oracle (rows {len(SAMPLE_ROWS)}): "
          + ", ".join(f"{NAMES.get(k, 'vector')}={v:.3e}" for k, v in errs.items()))
    for compute, e in errs.items():
        assert np.isfinite(e) and e <= bound[compute], (compute, e)
    # every row: the split paths against the vector path (worst row), within the sum of both bounds
    vec = outs[ff.COMPUTE_VECTOR]
    vnorm = torch.linalg.vector_norm(vec, dim=1)
    for compute in (ff.COMPUTE_FP16X3, ff.COMPUTE_BF16X3):
        d = torch.linalg.vector_norm(outs[compute] - vec, dim=1) / vnorm
        worst = float(d.max())
        print(f"  {NAMES[compute]} worst row vs vector path: {worst:.3e} (row {int(d.argmax())})")
        assert worst == worst and worst <= bound[compute] + bound[ff.COMPUTE_VECTOR]


# (prefetch depth, waves per SIMD) of the kernel variants, tuning variant = index + 1
MFMA_VARIANTS = [(2, 2), (3, 2), (1, 3), (2, 1), (4, 1)]


@pytest.mark.parametrize("compute", [ff.COMPUTE_FP16X3, ff.COMPUTE_BF16X3], ids=["fp16x3", "bf16x3"])
@pytest.mark.parametrize("direction", [ff.FORWARD, ff.INVERSE], ids=["fwd", "inv"])
@pytest.mark.parametrize("variant", range(1, len(MFMA_VARIANTS) + 1), ids=[f"pf{p}_wpe{w}" for p, w in MFMA_VARIANTS])
def test_wave_loop_and_prefetch_with_few_groups(gpu, compute, direction, variant):
    """Two workgroups (8 waves) for 203 transforms: every wave runs its loop
    25 or 26 times, each iteration on a transform an earlier one prefetched
    (a ring of PF transforms in flight; 203 leaves a partial last round)."""
    batch = 203
    x = complex_input(ora.accuracy_seed(int(ff.Type.C2C), N, batch, 10), (batch, N), np.complex64, "normal")
    ff.check("flagfftxSetTuning", ff.set_tuning(units=2, variant=variant))
    try:
        y, desc = _run(gpu, x, direction, compute)
    finally:
        ff.set_tuning()
    pf, wpe = MFMA_VARIANTS[variant - 1]
    assert "groups=2" in desc and f"{pf},{wpe}>" in desc, desc
    ref = ora.fft(x.astype(np.complex128), sign=direction)
    l2, _ = ora.rel_l2(y, ref, N)
    print(f"{NAMES[compute]} PF={pf} WPE={wpe}, 2 groups x 4 waves, batch {batch}: rel_l2={l2:.3e}")
    assert np.isfinite(l2) and l2 <= BOUND[compute]
