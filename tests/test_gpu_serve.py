"""Serving loop (flagfftxServer*): continuous batching, admission, alone == in the mix (GPU).

Every request's output is compared three ways:
  * bit-identical between the request served alone and served inside a mix
    of other lengths, types and precisions (same kernels, same arguments);
  * against the FP64 oracle within the stated bounds (FP32 2e-6 log2 N,
    FP64 1e-14 log2 N, relative L2 per transform);
  * PLAIN mode (per-request plan launches) against the same oracle.
"""
import math

import numpy as np
import pytest

from flagfft_amd import flagfft as ff
from oracle import oracle as ora

from .helpers import complex_input, d2h, dev_tensor, real_input, sync

pytestmark = pytest.mark.gpu

T = ff.Type
# (type, n, batch, direction): single-pass (queued) and multi-pass requests
MIX = [
    (T.C2C, 64, 5, ff.FORWARD), (T.C2C, 97, 3, ff.INVERSE), (T.C2C, 1000, 2, ff.FORWARD),
    (T.C2C, 4096, 4, ff.FORWARD), (T.C2C, 1009, 2, ff.FORWARD), (T.C2C, 2 * 3 * 5 * 7 * 11, 3, ff.INVERSE),
    (T.C2C, 20000, 2, ff.FORWARD), (T.C2C, 65536, 1, ff.FORWARD), (T.C2C, 10007, 2, ff.FORWARD),
    (T.R2C, 6000, 3, 0), (T.R2C, 256, 7, 0), (T.R2C, 999, 2, 0), (T.C2R, 512, 4, 0), (T.C2R, 6000, 2, 0),
    (T.Z2Z, 128, 3, ff.FORWARD), (T.Z2Z, 1009, 2, ff.INVERSE), (T.Z2Z, 10935, 1, ff.FORWARD),
    (T.Z2Z, 3000, 2, ff.FORWARD), (T.D2Z, 1000, 2, 0), (T.Z2D, 2048, 2, 0),
]
DBL = {T.Z2Z, T.D2Z, T.Z2D}


def _inputs(dev):
    xs, hosts = [], []
    for i, (t, n, b, _) in enumerate(MIX):
        dbl = t in DBL
        if t in (T.C2C, T.Z2Z):
            x = complex_input(500 + i, (b, n), np.complex128 if dbl else np.complex64, "normal")
        elif t in (T.R2C, T.D2Z):
            x = real_input(500 + i, (b, n), np.float64 if dbl else np.float32)
        else:
            full = ora.rfft(real_input(500 + i, (b, n), np.float64))
            x = full.astype(np.complex128 if dbl else np.complex64)
        hosts.append(x)
        xs.append(dev_tensor(x, dev))
    sync(dev)
    return xs, hosts


def _out_spec(t, n, b):
    dbl = t in DBL
    if t in (T.C2C, T.Z2Z):
        return (b, n), np.complex128 if dbl else np.complex64
    if t in (T.R2C, T.D2Z):
        return (b, n // 2 + 1), np.complex128 if dbl else np.complex64
    return (b, n), np.float64 if dbl else np.float32


def _oracle(t, n, x, d):
    x = np.asarray(x)
    if t in (T.C2C, T.Z2Z):
        return ora.fft(x.astype(np.complex128), sign=d)
    if t in (T.R2C, T.D2Z):
        return ora.rfft(x.astype(np.float64))
    return ora.irfft(x.astype(np.complex128), n)


def _bound(t, n):
    return (1e-14 if t in DBL else 2e-6) * max(math.log2(n), 1.0)


def _collect(srv, tickets):
    outs = []
    for tk, (t, n, b, _) in zip(tickets, MIX):
        ptr, lat = srv.wait(tk)
        assert lat >= 0
        shape, dt = _out_spec(t, n, b)
        outs.append(d2h(ptr, shape, dt))
        srv.release(tk)
    return outs


def test_each_result_is_the_same_alone_and_in_the_mix(gpu):
    xs, hosts = _inputs(gpu)
    with ff.Server(1 << 30, max_inflight=3) as srv:
        alone = []
        for x, (t, n, b, d) in zip(xs, MIX):
            tk = srv.submit(t, n, b, x.data_ptr(), d)
            ptr, _ = srv.wait(tk)
            shape, dt = _out_spec(t, n, b)
            alone.append(d2h(ptr, shape, dt))
            srv.release(tk)
        order = np.random.default_rng(3).permutation(len(MIX))
        tickets = [None] * len(MIX)
        for i in order:
            t, n, b, d = MIX[i]
            tickets[i] = srv.submit(t, n, b, xs[i].data_ptr(), d)
        mixed = _collect(srv, tickets)
        st = srv.stats()
    assert st["queue_launches"] >= 1 and st["queued_requests"] >= 10, st
    for (t, n, b, d), a, m, x in zip(MIX, alone, mixed, hosts):
        assert np.array_equal(a, m), f"{t.name} n={n}: result differs between alone and in the mix"
        l2, _ = ora.rel_l2(m, _oracle(t, n, x, d), m.shape[-1])
        assert l2 <= _bound(t, n), f"{t.name} n={n} rel_l2={l2:.3e}"


def test_admission_by_hbm_budget(gpu):
    """A budget that holds only a few requests at once still serves all of them, in bounds."""
    xs, hosts = _inputs(gpu)
    need = max(np.prod(_out_spec(t, n, b)[0]) * np.dtype(_out_spec(t, n, b)[1]).itemsize for t, n, b, _ in MIX)
    budget = int(need * 6)  # outputs and workspaces of a handful of requests
    with ff.Server(budget, max_inflight=2) as srv:
        tickets = [srv.submit(t, n, b, x.data_ptr(), d) for x, (t, n, b, d) in zip(xs, MIX)]
        outs = _collect(srv, tickets)
        st = srv.stats()
    assert st["completed"] == len(MIX)
    assert st["arena_peak_bytes"] <= budget
    for (t, n, b, d), y, x in zip(MIX, outs, hosts):
        l2, _ = ora.rel_l2(y, _oracle(t, n, x, d), y.shape[-1])
        assert l2 <= _bound(t, n), f"{t.name} n={n} rel_l2={l2:.3e}"


def test_plain_mode_matches_the_oracle(gpu):
    xs, hosts = _inputs(gpu)
    with ff.Server(1 << 30, mode=ff.SERVE_PLAIN) as srv:
        tickets = [srv.submit(t, n, b, x.data_ptr(), d) for x, (t, n, b, d) in zip(xs, MIX)]
        outs = _collect(srv, tickets)
        st = srv.stats()
    assert st["queue_launches"] == 0 and st["completed"] == len(MIX)
    assert st["transforms"] == sum(b for _, _, b, _ in MIX)
    assert 0 <= st["p50_ms"] <= st["p99_ms"] <= st["max_ms"]
    for (t, n, b, d), y, x in zip(MIX, outs, hosts):
        l2, _ = ora.rel_l2(y, _oracle(t, n, x, d), y.shape[-1])
        assert l2 <= _bound(t, n), f"{t.name} n={n} rel_l2={l2:.3e}"


def test_request_larger_than_the_budget_is_refused(gpu):
    x = dev_tensor(np.zeros((4, 4096), np.complex64), gpu)
    with ff.Server(64 << 10, max_inflight=1) as srv:
        with pytest.raises(ff.FlagFFTError) as e:
            srv.submit(T.C2C, 4096, 4, x.data_ptr(), ff.FORWARD)
        assert e.value.code == ff.Result.ALLOC_FAILED


def test_request_past_2_31_elements_is_refused(gpu):
    """nx * batch above 2^31 - 1 is refused before anything is allocated or launched."""
    x = dev_tensor(np.zeros((1, 64), np.complex64), gpu)
    with ff.Server(1 << 20, max_inflight=1) as srv:
        with pytest.raises(ff.FlagFFTError) as e:
            srv.submit(T.C2C, 4096, (1 << 19) + 1, x.data_ptr(), ff.FORWARD)
        assert e.value.code == ff.Result.INVALID_SIZE
        assert srv.stats()["completed"] == 0


# ---- the serving configuration's request stream ------------------------------
def _trace(nreq: int):
    """Requests of the serving configuration (length log-uniform in [64, 2^20],
    20% primes, batch uniform in [1, 512], 70% C2C f32 / 20% R2C f32 / 10% Z2Z
    f64), the same generator the benchmark's serving leg uses, plus one
    2^20 x 256 C2C request and one prime-length request so that both kinds
    are always present."""
    from bench import _is_prime, serving_trace

    tm = {"C2C": T.C2C, "R2C": T.R2C, "Z2Z": T.Z2Z}
    reqs = [(tm[t], n, b) for t, n, b in serving_trace(nreq, 2024)]
    reqs.insert(nreq // 3, (T.C2C, 1 << 20, 256))
    reqs.insert(2 * nreq // 3, (T.C2C, 65537, 97))
    assert any(_is_prime(n) for _, n, _ in reqs)
    return reqs


def _sample_rows(b: int):
    return sorted({0, b // 2, b - 1})


def test_serving_trace_poisson_alone_and_in_the_mix(gpu):
    """The serving configuration's stream at reduced length, with Poisson
    arrivals: every request's output is bit-identical served alone and in the
    mix, and its first, middle and last transforms match the FP64 oracle."""
    import time

    import torch

    from .helpers import hip_runtime

    reqs = _trace(30)
    # one input buffer per type, sized for its largest request; every request reads its prefix
    count = {T.C2C: 2, T.R2C: 1, T.Z2Z: 2}
    dtype = {T.C2C: torch.float32, T.R2C: torch.float32, T.Z2Z: torch.float64}
    need = {t: max([count[t] * n * b for tt, n, b in reqs if tt == t] or [2]) for t in count}
    g = torch.Generator(device=gpu)
    g.manual_seed(77)
    inputs = {t: torch.randn(need[t], dtype=dtype[t], device=gpu, generator=g) for t in count}
    host_in = {t: v.cpu().numpy() for t, v in inputs.items()}
    sync(gpu)

    def out_bytes(t, n, b):
        shape, dt = _out_spec(t, n, b)
        return int(np.prod(shape)) * np.dtype(dt).itemsize

    def keep(ptr, t, n, b):
        dst = torch.empty(out_bytes(t, n, b), dtype=torch.uint8, device=gpu)
        assert hip_runtime().hipMemcpy(dst.data_ptr(), ptr, dst.numel(), 3) == 0  # device to device
        return dst

    with ff.Server(32 << 30, max_inflight=3) as srv:
        alone = []
        for t, n, b in reqs:
            tk = srv.submit(t, n, b, inputs[t].data_ptr(), ff.FORWARD)
            ptr, _ = srv.wait(tk)
            alone.append(keep(ptr, t, n, b))
            srv.release(tk)
        srv.reset_stats()
        rng = np.random.default_rng(5)
        gaps = rng.exponential(0.4e-3, len(reqs))  # Poisson arrivals, 2500 requests/s offered
        tickets = []
        t_next = time.perf_counter()
        for (t, n, b), gap in zip(reqs, gaps):
            t_next += gap
            while time.perf_counter() < t_next:
                pass
            tickets.append(srv.submit(t, n, b, inputs[t].data_ptr(), ff.FORWARD))
        mixed = []
        for tk, (t, n, b) in zip(tickets, reqs):
            ptr, lat = srv.wait(tk)
            assert lat >= 0
            mixed.append(keep(ptr, t, n, b))
            srv.release(tk)
        st = srv.stats()
    assert st["completed"] == len(reqs)
    print(f"serving trace: {len(reqs)} requests, {st['transforms']} transforms, queue launches "
          f"{st['queue_launches']}, queued requests {st['queued_requests']}, p50 {st['p50_ms']:.3f} ms, "
          f"p99 {st['p99_ms']:.3f} ms")
    for (t, n, b), a, m in zip(reqs, alone, mixed):
        assert torch.equal(a, m), f"{t.name} n={n} batch={b}: result differs between alone and in the mix"
        rows = _sample_rows(b)
        shape, dt = _out_spec(t, n, b)
        y = m.view(b, -1)[rows].cpu().numpy().view(dt).reshape(len(rows), shape[-1])
        flat = host_in[t]
        if t == T.C2C:
            x = flat[: 2 * n * b].view(np.complex64).reshape(b, n)[rows]
        elif t == T.Z2Z:
            x = flat[: 2 * n * b].view(np.complex128).reshape(b, n)[rows]
        else:
            x = flat[: n * b].reshape(b, n)[rows]
        l2, _ = ora.rel_l2(y, _oracle(t, n, x, ff.FORWARD), y.shape[-1])
        assert l2 <= _bound(t, n), f"{t.name} n={n} batch={b} rel_l2={l2:.3e}"


# Multi-pass requests whose every pass has a persistent-queue form: register-
# resident passes of 256 threads (power-of-two super-radices), the real split
# pass, FP64 passes of 64 and 256 points.  Each request's passes run as
# queue launches of the batch (step s of every request in the s-th launches).
MULTI = [
    (T.C2C, 1 << 16, 3, ff.FORWARD), (T.C2C, 1 << 20, 2, ff.INVERSE), (T.C2C, 1 << 18, 5, ff.FORWARD),
    (T.C2C, 1 << 14, 9, ff.INVERSE), (T.R2C, 1 << 17, 3, 0), (T.Z2Z, 1 << 16, 2, ff.FORWARD),
    (T.C2C, 1 << 13, 7, ff.FORWARD),
]


def test_multi_pass_requests_run_in_queue_launches(gpu):
    """Every MULTI request goes through the persistent queue (one launch per
    pass and kernel class), gives the same bits alone and in the mix, and
    matches the FP64 oracle."""
    import torch

    xs, hosts = [], []
    for i, (t, n, b, _) in enumerate(MULTI):
        dbl = t in DBL
        if t in (T.C2C, T.Z2Z):
            x = complex_input(900 + i, (b, n), np.complex128 if dbl else np.complex64, "normal")
        else:
            x = real_input(900 + i, (b, n), np.float32)
        hosts.append(x)
        xs.append(dev_tensor(x, gpu))
    sync(gpu)
    with ff.Server(4 << 30, max_inflight=3) as srv:
        alone = []
        for x, (t, n, b, d) in zip(xs, MULTI):
            tk = srv.submit(t, n, b, x.data_ptr(), d)
            ptr, _ = srv.wait(tk)
            shape, dt = _out_spec(t, n, b)
            alone.append(d2h(ptr, shape, dt))
            srv.release(tk)
        st_alone = srv.stats()
        srv.reset_stats()
        tickets = [srv.submit(t, n, b, x.data_ptr(), d) for x, (t, n, b, d) in zip(xs, MULTI)]
        mixed = []
        for tk, (t, n, b, _) in zip(tickets, MULTI):
            ptr, _ = srv.wait(tk)
            shape, dt = _out_spec(t, n, b)
            mixed.append(d2h(ptr, shape, dt))
            srv.release(tk)
        st = srv.stats()
    print(f"multi-pass alone: {st_alone['queued_requests']} queued, {st_alone['queue_launches']} launches; "
          f"mix: {st['queued_requests']} queued, {st['queue_launches']} launches")
    assert st_alone["queued_requests"] == len(MULTI) and st["queued_requests"] == len(MULTI), (st_alone, st)
    # alone, each request takes one launch per pass at least: more launches than requests
    assert st_alone["queue_launches"] > len(MULTI)
    for (t, n, b, d), a, m, x in zip(MULTI, alone, mixed, hosts):
        assert np.array_equal(a, m), f"{t.name} n={n}: result differs between alone and in the mix"
        l2, _ = ora.rel_l2(m, _oracle(t, n, x, d), m.shape[-1])
        assert l2 <= _bound(t, n), f"{t.name} n={n} rel_l2={l2:.3e}"


def test_small_requests_pass_one_that_does_not_fit_yet(gpu):
    """BATCHED admission: while a request waits for arena memory (an earlier
    result is still held by the client), smaller requests behind it are
    admitted and complete; the waiting one runs once the memory is released."""
    import time

    n, b = 65536, 16  # two-pass chain: 8 MiB of output and 8 MiB of workspace
    big = complex_input(901, (b, n), np.complex64, "normal")
    small = [complex_input(902 + i, (5, 64), np.complex64, "normal") for i in range(4)]
    big_d = dev_tensor(big, gpu)
    small_d = [dev_tensor(x, gpu) for x in small]
    sync(gpu)

    def poll_until(srv, tk, seconds):
        t0 = time.time()
        while time.time() - t0 < seconds:
            if srv.poll(tk)[0]:
                return True
            time.sleep(1e-3)
        return False

    with ff.Server(20 << 20, max_inflight=2) as srv:
        holder = srv.submit(T.C2C, n, b, big_d.data_ptr(), ff.FORWARD)
        assert poll_until(srv, holder, 30.0)  # done; its output stays held
        waiting = srv.submit(T.C2C, n, b, big_d.data_ptr(), ff.FORWARD)  # 8 + 16 MiB > 20 MiB
        smalls = [srv.submit(T.C2C, 64, 5, x.data_ptr(), ff.FORWARD) for x in small_d]
        for tk, x in zip(smalls, small):
            assert poll_until(srv, tk, 30.0), "a small request was held behind one that does not fit"
            ptr, _ = srv.wait(tk)
            y = d2h(ptr, (5, 64), np.complex64)
            srv.release(tk)
            l2, _ = ora.rel_l2(y, ora.fft(x.astype(np.complex128)), 64)
            assert l2 <= _bound(T.C2C, 64)
        assert not srv.poll(waiting)[0], "the waiting request ran without its memory"
        srv.release(holder)
        ptr, _ = srv.wait(waiting)
        y = d2h(ptr, (b, n), np.complex64)
        srv.release(waiting)
    rows = [0, b - 1]
    l2, _ = ora.rel_l2(y[rows], ora.fft(big[rows].astype(np.complex128)), n)
    assert l2 <= _bound(T.C2C, n)
