"""Jobs past 2^31 complex elements (GPU): the single-pass leaf (N = 4096,
batch 2^19 + 3) and the two-pass Stockham chain (N = 2^24, batch 129, with
its workspace of the same size) against the oracle on sampled sequences on
both sides of the 2^29-element (4 GiB) and 2^31-element boundaries, where a
32-bit byte or element offset would wrap.  About 35 GB (leaf) and 52 GB
(chain) of device memory.
Bounds: FP32 2e-6 log2 N relative L2 per sampled transform.
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
