"""Graph mode (flagfftxSetGraphMode): every replay equals the direct launches (GPU).

The graph holds the same kernels with the same arguments, so results are
bit-identical.  Replays alternate buffers (the nodes touching the caller's
buffers are patched), with unrelated eager work and synchronisations between
them, and every replay is checked.
"""
import numpy as np
import pytest

from flagfft_amd import flagfft as ff

from .helpers import complex_input, dev_tensor, host, real_input, stream_of, sync

pytestmark = pytest.mark.gpu

CASES = [
    ("c2c", 4096, 8, ff.Type.C2C),          # one specialised kernel
    ("c2c", 1 << 20, 2, ff.Type.C2C),       # two super-stages
    ("c2c", 10007, 3, ff.Type.C2C),         # Bluestein, four passes
    ("c2c", 10935, 2, ff.Type.Z2Z),
    ("r2c", 6000, 4, ff.Type.R2C),
    ("c2r", 6000, 4, ff.Type.C2R),
    ("r2c", 2 * 10007, 2, ff.Type.D2Z),
]


def _exec(p, kind, t, i, o):
    if kind == "c2c":
        (p.z2z if t == ff.Type.Z2Z else p.c2c)(i.data_ptr(), o.data_ptr(), ff.FORWARD)
    elif kind == "r2c":
        (p.d2z if t == ff.Type.D2Z else p.r2c)(i.data_ptr(), o.data_ptr())
    else:
        p.c2r(i.data_ptr(), o.data_ptr())


@pytest.mark.parametrize("kind,n,batch,t", CASES, ids=[f"{c[0]}-{c[1]}-{c[3].name}" for c in CASES])
def test_graph_replays_match_direct_launches(gpu, kind, n, batch, t):
    import torch

    dbl = t in (ff.Type.Z2Z, ff.Type.D2Z)
    cdt = np.complex128 if dbl else np.complex64
    rdt = np.float64 if dbl else np.float32
    h = n // 2 + 1
    inputs = []
    for v in range(3):
        if kind == "c2c":
            x = complex_input(1000 + v, (batch, n), cdt, "normal")
        elif kind == "r2c":
            x = real_input(1000 + v, (batch, n), rdt)
        else:
            x = complex_input(1000 + v, (batch, h), cdt, "normal")
        inputs.append(dev_tensor(x, gpu))
    out_shape = (batch, n) if kind == "c2c" else ((batch, h) if kind == "r2c" else (batch, n))
    out_dt = torch.complex128 if dbl else torch.complex64
    if kind == "c2r":
        out_dt = torch.float32
    with ff.Plan.one_d(n, t, batch) as p:
        p.set_stream(stream_of(gpu))
        direct = []
        for x in inputs:
            y = torch.zeros(out_shape, dtype=out_dt, device=gpu)
            _exec(p, kind, t, x, y)
            sync(gpu)
            direct.append(host(y))
        p.set_graph_mode(True)
        for rep in range(2):
            for v, x in enumerate(inputs):
                y = torch.full(out_shape, 7.0, dtype=out_dt, device=gpu)
                _exec(p, kind, t, x, y)
                sync(gpu)
                junk = torch.randn(1 << 16, device=gpu).sum()  # unrelated eager work between replays
                sync(gpu)
                assert np.array_equal(host(y), direct[v]), f"replay {rep} input {v} differs ({junk.item():.1f})"
