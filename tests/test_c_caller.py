"""The library called from C through include/flagfft.h only (tests/c_caller/abi_caller.c,
built by __graft_entry__.build()): contract checks, then C2C and R2C/C2R roundtrips on the GPU."""
import os
import subprocess

import pytest

EXE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "c_caller", "abi_caller")


def _run(*args):
    assert os.path.exists(EXE), "tests/c_caller/abi_caller is not built; run __graft_entry__.build()"
    r = subprocess.run([EXE, *args], capture_output=True, text=True, timeout=120)
    return r.returncode, r.stdout + r.stderr


def test_c_caller_without_device():
    import torch

    if torch.cuda.is_available():
        pytest.skip("a GPU is visible")
    rc, out = _run("--no-device")
    assert rc == 0 and "abi caller ok" in out, out


@pytest.mark.gpu
def test_c_caller_on_the_gpu(gpu):
    rc, out = _run()
    assert rc == 0 and "abi caller ok" in out, out
