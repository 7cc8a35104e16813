"""flagfft_amd — MI355X-native (gfx950) FFT library behind FlagFFT's C API.

The product is the C-ABI shared library ``flagfft_amd/libflagfft.so``
(declared in include/flagfft.h and include/flagfft_ext.h).  This package only
holds its sources, its build recipe and a ctypes mirror of the C API
(``flagfft_amd.flagfft``) used by tests and benchmarks.
"""
from . import flagfft  # noqa: F401
from .flagfft import FORWARD, INVERSE, FlagFFTError, Plan, Result, Type  # noqa: F401

__all__ = ["flagfft", "Plan", "Result", "Type", "FORWARD", "INVERSE", "FlagFFTError"]
