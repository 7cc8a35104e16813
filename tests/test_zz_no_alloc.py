"""flagfftxExecC2CNoAlloc against flagfftExecC2C (GPU, through the C ABI).

The allocation-free entry point runs a plan's kernels with their workspaces in
the library's fixed per-device array instead of the plan's own.  Its results
must be bit-identical to flagfftExecC2C on the same plan and input, on two
calls in a row whose second has the smaller input (the array keeps the first
call's contents), and within the FP32 bounds of the FP64 oracle.  Plans whose
workspaces exceed the array are refused with FLAGFFT_INVALID_SIZE.
"""
import numpy as np
import pytest

from flagfft_amd import flagfft as ff
from oracle import oracle as ora

from .helpers import check_accuracy, complex_input, dev_tensor, host, stream_of, sync

pytestmark = pytest.mark.gpu

# (n, batch): a two-pass chain with a 16 MiB workspace, then smaller plans:
# a two-pass chain of 1.5 MiB, the two-halves LDS Bluestein (park workspace),
# a three-launch Bluestein chain, and a single-pass leaf (no workspace)
CASES = [(1 << 20, 2), (1 << 16, 3), (12011, 2), (40009, 2), (4096, 5)]


def _both(dev, n, batch, direction, seed):
    import torch

    x = complex_input(seed, (batch, n), np.complex64, "normal")
    xd = dev_tensor(x, dev)
    ya = torch.empty_like(xd)
    yb = torch.empty_like(xd)
    with ff.Plan.one_d(n, ff.Type.C2C, batch) as p:
        p.set_stream(stream_of(dev))
        ff.check("flagfftxExecC2CNoAlloc", ff.exec_c2c_noalloc(p.handle, xd.data_ptr(), ya.data_ptr(), direction))
        ff.check("flagfftExecC2C", ff.exec_c2c(p.handle, xd.data_ptr(), yb.data_ptr(), direction))
        sync(dev)
        desc = p.description()
    return x, host(ya), host(yb), desc


@pytest.mark.parametrize("direction", [ff.FORWARD, ff.INVERSE], ids=["fwd", "inv"])
def test_matches_exec_c2c_on_calls_in_a_row(gpu, direction):
    for k, (n, batch) in enumerate(CASES):
        print(f"no-alloc C2C n={n} batch={batch} dir={direction}", flush=True)
        x, ya, yb, desc = _both(gpu, n, batch, direction, 0xA110C + 31 * k + direction)
        assert np.array_equal(ya.view(np.uint32), yb.view(np.uint32)), f"n={n}: differs from flagfftExecC2C\n{desc}"
        check_accuracy(ya, ora.fft(x.astype(np.complex128), sign=direction), ff.Type.C2C, n, row_len=n,
                       what=f"no-alloc n={n}")


def test_in_place_matches(gpu):
    """In place (input = output), where split leaves use their in-place executable."""
    import torch

    n, batch = 10935, 3
    x = complex_input(0xA11C, (batch, n), np.complex64, "normal")
    a = dev_tensor(x, gpu)
    b = a.clone()
    with ff.Plan.one_d(n, ff.Type.C2C, batch) as p:
        p.set_stream(stream_of(gpu))
        ff.check("flagfftxExecC2CNoAlloc", ff.exec_c2c_noalloc(p.handle, a.data_ptr(), a.data_ptr(), ff.FORWARD))
        ff.check("flagfftExecC2C", ff.exec_c2c(p.handle, b.data_ptr(), b.data_ptr(), ff.FORWARD))
        sync(gpu)
    assert torch.equal(a, b)


def test_refusals(gpu):
    import torch

    x = torch.zeros((1, 1 << 24), dtype=torch.complex64, device=gpu)
    y = torch.empty_like(x)
    # a multi-pass 2^24 chain needs at least one 128 MiB workspace: above the 64 MiB array
    with ff.Plan.one_d(1 << 24, ff.Type.C2C, 1) as p:
        assert ff.exec_c2c_noalloc(p.handle, x.data_ptr(), y.data_ptr(), ff.FORWARD) == ff.Result.INVALID_SIZE
        assert ff.exec_c2c_noalloc(p.handle, x.data_ptr(), y.data_ptr(), 0) == ff.Result.INVALID_VALUE
        assert ff.exec_c2c_noalloc(p.handle, None, y.data_ptr(), ff.FORWARD) == ff.Result.INVALID_VALUE
    with ff.Plan.one_d(256, ff.Type.Z2Z, 1) as p:
        assert ff.exec_c2c_noalloc(p.handle, x.data_ptr(), y.data_ptr(), ff.FORWARD) == ff.Result.INVALID_TYPE
    r, h = ff.plan1d(256, ff.Type.C2C, 1)
    ff.check("flagfftPlan1d", r)
    ff.check("flagfftDestroy", ff.destroy(h))
    assert ff.exec_c2c_noalloc(h, x.data_ptr(), y.data_ptr(), ff.FORWARD) == ff.Result.INVALID_PLAN
    sync(gpu)
