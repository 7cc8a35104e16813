"""Shared helpers for the parity tests: inputs, tolerances, device round trips.

Tolerances (both must hold, the tighter one decides):
  * the reference's accuracy model rel <= C_class * u * work(N)
    (ctest/flagfft_test.h:274-320; constants frozen at 2x cuFFT's worst case);
  * the stated bounds FP32 rel_l2 <= 2e-6 * log2 N, FP64 <= 1e-14 * log2 N.
"""
from __future__ import annotations

import math

import numpy as np

from flagfft_amd import flagfft as ff
from oracle import oracle as ora

# ctest/flagfft_test.h:285-297 (rel_l2, rel_linf)
CLASS_CONSTANTS = {
    "complex": (1.2419386546059821, 1.9343969087678796),
    "real_forward": (1.234681000407627, 1.8260558195934091),
    "real_inverse": (0.97722970418819066, 1.372182697342486),
}

DOUBLE_TYPES = {ff.Type.Z2Z, ff.Type.D2Z, ff.Type.Z2D}


def tclass(t) -> str:
    if t in (ff.Type.C2C, ff.Type.Z2Z):
        return "complex"
    if t in (ff.Type.R2C, ff.Type.D2Z):
        return "real_forward"
    return "real_inverse"


def unit_roundoff(t) -> float:
    return (np.finfo(np.float64).eps if t in DOUBLE_TYPES else np.finfo(np.float32).eps) / 2.0


def ref_limits(t, n: int):
    c2, cinf = CLASS_CONSTANTS[tclass(t)]
    s = unit_roundoff(t) * ora.work_factor(n)
    return c2 * s, cinf * s


def stated_bound(t, n: int) -> float:
    lg = max(math.log2(n), 1.0)
    return (1e-14 if t in DOUBLE_TYPES else 2e-6) * lg


def check_accuracy(value, ref, t, n: int, row_len: int | None = None, factor: float = 1.0, what: str = ""):
    l2, linf = ora.rel_l2(value, ref, row_len)
    lim2, liminf = ref_limits(t, n)
    bound = stated_bound(t, n)
    msg = (f"{what} type={ff.Type(t).name} N={n} rel_l2={l2:.3e} (ref-model {factor * lim2:.3e}, stated {factor * bound:.3e})"
           f" rel_linf={linf:.3e} (ref-model {factor * liminf:.3e})")
    assert np.isfinite(l2) and l2 <= factor * lim2 and l2 <= factor * bound and linf <= factor * liminf, msg
    return l2, linf


def dev_tensor(a: np.ndarray, dev):
    import torch

    return torch.from_numpy(np.ascontiguousarray(a)).to(dev)


def host(t) -> np.ndarray:
    return t.detach().cpu().numpy()


def stream_of(dev) -> int:
    import torch

    return torch.cuda.current_stream(dev).cuda_stream


def sync(dev):
    import torch

    torch.cuda.synchronize(dev)


def complex_input(seed: int, shape, dtype, kind: str = "uniform", scale: float = 1.0) -> np.ndarray:
    if kind == "uniform":
        x = ora.complex_signed_unit(seed, shape)
    else:
        x = ora.complex_normal(seed, shape)
    return (x * scale).astype(dtype)


def real_input(seed: int, shape, dtype, kind: str = "uniform", scale: float = 1.0) -> np.ndarray:
    count = int(np.prod(shape))
    x = ora.signed_unit(seed, count) if kind == "uniform" else ora.normal(seed, count)
    return (x.reshape(shape) * scale).astype(dtype)


def run_c2c(dev, x: np.ndarray, direction: int, n: int, batch: int, inplace: bool = False) -> np.ndarray:
    dbl = x.dtype == np.complex128
    t = ff.Type.Z2Z if dbl else ff.Type.C2C
    xd = dev_tensor(x, dev)
    yd = xd if inplace else xd.clone().zero_()
    with ff.Plan.one_d(n, t, batch) as p:
        p.set_stream(stream_of(dev))
        (p.z2z if dbl else p.c2c)(xd.data_ptr(), yd.data_ptr(), direction)
        sync(dev)
    return host(yd)


_HIP = None


def hip_runtime():
    """The HIP runtime library this process already loaded (torch's and the library's are one)."""
    global _HIP
    if _HIP is None:
        import ctypes
        import os
        import re

        paths = sorted(set(re.findall(r"\S*libamdhip64\.so\S*", open("/proc/self/maps").read())))
        assert paths, "HIP runtime not loaded"
        _HIP = ctypes.CDLL(paths[0], mode=os.RTLD_NOLOAD | os.RTLD_GLOBAL)
        _HIP.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int]
        _HIP.hipMemcpy.restype = ctypes.c_int
    return _HIP


def d2h(ptr: int, shape, dtype) -> np.ndarray:
    """Copy a raw device allocation into a new host array (synchronous)."""
    out = np.empty(shape, dtype=dtype)
    rc = hip_runtime().hipMemcpy(out.ctypes.data, ptr, out.nbytes, 2)  # hipMemcpyDeviceToHost
    assert rc == 0, f"hipMemcpy failed ({rc})"
    return out
