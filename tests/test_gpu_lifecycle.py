"""Plan lifetime and threads (GPU), after the reference's threading contract:
one mutex per plan serialises Exec, SetStream and GetPlanDescription, and
different plans may be used concurrently (src/exec/c_api_exec.cpp:169-236,
src/plan/c_api_internal.hpp:5-7).

* Destroy racing host-side calls on the same handle: every call returns
  SUCCESS or INVALID_PLAN, none touches a freed plan, and once Destroy has
  returned the handle is INVALID_PLAN.  Only calls that launch nothing race
  the Destroy, so no kernel is in flight while the plan's memory is freed.
* Two plans executed from two host threads at once, each on its own stream
  and buffers, give bit-identical results to the same calls made one after
  the other on one thread.
"""
import ctypes
import threading
import time

import numpy as np
import pytest

from flagfft_amd import flagfft as ff

from .helpers import complex_input, dev_tensor, host

pytestmark = pytest.mark.gpu


def test_destroy_races_host_calls(gpu):
    R = ff.Result
    for rep in range(4):
        p = ff.Plan.one_d(4096, ff.Type.C2C, 4)
        h = p.handle
        seen = []
        stop = threading.Event()

        def hammer():
            codes = set()
            for _ in range(200000):
                codes.add(int(ff.set_stream(h, 0)))
                codes.add(int(ff.plan_stats(h)[0]))
                if R.INVALID_PLAN in codes or stop.is_set():
                    break
            seen.append(codes)

        threads = [threading.Thread(target=hammer) for _ in range(4)]
        for t in threads:
            t.start()
        time.sleep(0.002 * (rep + 1))
        assert ff.destroy(h) == R.SUCCESS
        p.handle = None
        stop.set()
        for t in threads:
            t.join(timeout=60)
            assert not t.is_alive()
        assert len(seen) == 4
        for codes in seen:
            assert codes <= {int(R.SUCCESS), int(R.INVALID_PLAN)}, codes
        assert ff.set_stream(h, 0) == R.INVALID_PLAN
        assert ff.plan_stats(h)[0] == R.INVALID_PLAN
        assert ff.destroy(h) == R.INVALID_PLAN


def test_two_plans_from_two_threads(gpu):
    import torch

    cases = [(4096, 64), (10007, 16)]
    xs = [complex_input(n + b, (b, n), np.complex64) for n, b in cases]
    plans = [ff.Plan.one_d(n, ff.Type.C2C, b) for n, b in cases]
    streams = [torch.cuda.Stream(device=gpu) for _ in cases]
    try:
        for p, s in zip(plans, streams):
            p.set_stream(s.cuda_stream)
        ins = [dev_tensor(x, gpu) for x in xs]
        serial = []
        for p, xd in zip(plans, ins):
            yd = torch.zeros_like(xd)
            assert ff.exec_c2c(p.handle, xd.data_ptr(), yd.data_ptr(), ff.FORWARD) == ff.Result.SUCCESS
            torch.cuda.synchronize(gpu)
            serial.append(host(yd))
        outs = [[torch.zeros_like(xd) for _ in range(8)] for xd in ins]
        codes = [[], []]

        def run(i):
            for yd in outs[i]:
                codes[i].append(int(ff.exec_c2c(plans[i].handle, ins[i].data_ptr(), yd.data_ptr(), ff.FORWARD)))

        threads = [threading.Thread(target=run, args=(i,)) for i in range(2)]
        for t in threads:
            t.start()
        for t in threads:
            t.join(timeout=60)
            assert not t.is_alive()
        torch.cuda.synchronize(gpu)
        for i in range(2):
            assert codes[i] == [int(ff.Result.SUCCESS)] * 8
            for yd in outs[i]:
                assert np.array_equal(host(yd), serial[i]), f"N={cases[i][0]}"
    finally:
        torch.cuda.synchronize(gpu)
        for p in plans:
            p.destroy()


def test_destroyed_handle_is_never_reused(gpu):
    """A destroyed handle keeps returning FLAGFFT_INVALID_PLAN after new plans
    are created: its address is not handed out again."""
    import torch

    x = torch.zeros((2, 256), dtype=torch.complex64, device=gpu)
    r, old = ff.plan1d(256, ff.Type.C2C, 2)
    ff.check("flagfftPlan1d", r)
    stale = old.value
    ff.check("flagfftDestroy", ff.destroy(old))
    fresh = []
    try:
        for _ in range(64):
            r, h = ff.plan1d(256, ff.Type.C2C, 2)
            ff.check("flagfftPlan1d", r)
            fresh.append(h)
        assert stale not in {h.value for h in fresh}
        assert ff.exec_c2c(ctypes.c_void_p(stale), x.data_ptr(), x.data_ptr(), ff.FORWARD) == ff.Result.INVALID_PLAN
        assert ff.destroy(ctypes.c_void_p(stale)) == ff.Result.INVALID_PLAN
        assert ff.exec_c2c(fresh[-1], x.data_ptr(), x.data_ptr(), ff.FORWARD) == ff.Result.SUCCESS
        torch.cuda.synchronize(gpu)
    finally:
        for h in fresh:
            ff.destroy(h)
