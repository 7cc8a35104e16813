"""Benchmark of the FFT execution path behind include/flagfft.h on one MI355X.

Default workload (arbitrary-length 1D, one "step" = one pass of the hot path
over one batch of synthetic input):
  C2C f32 and Z2Z f64 forward at N=10007 (prime) and N=10935 = 3^7*5, batch 64, N(0,1);
  R2C->C2R f32 and D2Z->Z2D f64 at N=6000, batch 128, U[-1,1].
Every transform goes through the C ABI (libflagfft.so).
Inputs rotate over buffer sets larger than 2x the 256 MB Infinity Cache so
the timed region reads HBM.

value          algorithmic HBM GB/s of the whole step: sum over transforms of
               batch * (input bytes + output bytes), divided by wall time; the
               transform families (independent) are spread longest-first by
               their device time alone over 1 to 4 HIP streams (4 hardware
               queues per process), the count chosen by a short calibration
               region before the timed one; config.single_stream_gbps is the
               same step on one stream
roofline       dominant kernel (by total device time): algorithmic bytes per
               launch / mean launch duration from HIP events recorded on the
               plan's stream around every launch of a second timed region of
               the same K steps (the first, event-free region gives value);
               traffic = counter-measured HBM bytes per launch (committed
               rocprofv3 summary), traffic_gbps / traffic_frac = those bytes
               over the same duration
cpu_baseline   the oracle's fast FFT (oracle/oracle.c) on the host cores over
               a bounded sample of the same step

Usage: python bench.py [--gpus 1] [--steps K] [--warmup W] [--no-cpu] [--suite]
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "C2C/R2C effective HBM GB/s (% of peak); mixed-length serving transforms/s + p99"
HBM_PEAK_GBS = 8000.0
MALL_BYTES = 256 << 20


def log(*a):
    print(*a, file=sys.stderr, flush=True)


class Case:
    """One transform family of the workload: plans, rotating device buffers, one exec."""

    def __init__(self, name, kind, n, batch, dbl, dev, ff, torch, compute=0):
        self.name, self.kind, self.n, self.batch, self.dbl = name, kind, n, batch, dbl
        self.compute = compute
        self.ff, self.torch, self.dev = ff, torch, dev
        cdt = torch.complex128 if dbl else torch.complex64
        rdt = torch.float64 if dbl else torch.float32
        esz = 16 if dbl else 8
        T = ff.Type
        if isinstance(n, tuple):  # 2D / 3D (row-major, last axis contiguous), one transform per exec
            lead, nl = n[:-1], n[-1]
            h = nl // 2 + 1
            outer = 1
            for v in lead:
                outer *= v
            self.batch = 1

            def plan(t):
                return ff.Plan.two_d(n[0], n[1], t) if len(n) == 2 else ff.Plan.three_d(n[0], n[1], n[2], t, 1)

            if kind == "c2c":
                self.plans = [plan(T.Z2Z if dbl else T.C2C)]
                self.alg_bytes = outer * nl * esz * 2
                self.shapes = [(n, cdt), (n, cdt)]
            else:
                self.plans = [plan(T.D2Z if dbl else T.R2C), plan(T.Z2D if dbl else T.C2R)]
                self.alg_bytes = 2 * (outer * nl * esz // 2 + outer * h * esz)
                self.shapes = [(n, rdt), (lead + (h,), cdt), (n, rdt)]
            self.sets = []
            return
        h = n // 2 + 1
        if kind == "c2c":
            self.plans = [ff.Plan.one_d_compute(n, T.Z2Z if dbl else T.C2C, batch, compute)]
            self.alg_bytes = batch * n * esz * 2
            self.shapes = [((batch, n), cdt), ((batch, n), cdt)]
        else:  # real roundtrip R2C -> C2R
            self.plans = [ff.Plan.one_d(n, T.D2Z if dbl else T.R2C, batch),
                          ff.Plan.one_d(n, T.Z2D if dbl else T.C2R, batch)]
            self.alg_bytes = 2 * batch * (n * esz // 2 + h * esz)
            self.shapes = [((batch, n), rdt), ((batch, h), cdt), ((batch, n), rdt)]
        self.sets = []

    def footprint(self):
        return sum(int(np.prod(s)) * self.torch.empty((), dtype=d).element_size() for s, d in self.shapes)

    def allocate(self, nsets, seed):
        torch = self.torch
        g = torch.Generator(device=self.dev)
        g.manual_seed(seed)
        for _ in range(nsets):
            bufs = []
            for i, (shape, dt) in enumerate(self.shapes):
                if i == 0:
                    if dt.is_complex:
                        t = torch.randn(shape, dtype=dt, device=self.dev, generator=g)  # re, im iid N(0, 1/2)
                        t = t * (2.0 ** 0.5)                                             # -> iid N(0, 1)
                    else:
                        t = torch.rand(shape, dtype=dt, device=self.dev, generator=g) * 2 - 1  # U[-1, 1)
                else:
                    t = torch.empty(shape, dtype=dt, device=self.dev)
                bufs.append(t)
            self.sets.append(bufs)

    def bind(self, stream):
        for p in self.plans:
            p.set_stream(stream)

    def run(self, step):
        b = self.sets[step % len(self.sets)]
        p = self.plans
        if self.kind == "c2c":
            (p[0].z2z if self.dbl else p[0].c2c)(b[0].data_ptr(), b[1].data_ptr(), self.ff.FORWARD)
        else:
            (p[0].d2z if self.dbl else p[0].r2c)(b[0].data_ptr(), b[1].data_ptr())
            (p[1].z2d if self.dbl else p[1].c2r)(b[1].data_ptr(), b[2].data_ptr())

    def cpu_step(self, ora, nthreads, frac):
        """Oracle fast FFT over a fraction of the batch; returns (seconds, algorithmic bytes)."""
        nb = max(1, int(self.batch * frac))
        rng = np.random.default_rng(self.n)
        if self.kind == "c2c":
            x = (rng.standard_normal((nb, self.n)) + 1j * rng.standard_normal((nb, self.n)))
            x = x.astype(np.complex128 if self.dbl else np.complex64)
            t0 = time.perf_counter()
            (ora.fft if self.dbl else ora.fft_f32)(x, -1, nthreads)
            dt = time.perf_counter() - t0
        else:
            x = rng.uniform(-1, 1, (nb, self.n)).astype(np.float64 if self.dbl else np.float32)
            t0 = time.perf_counter()
            X = (ora.rfft if self.dbl else ora.rfft_f32)(x, nthreads)
            (ora.irfft if self.dbl else ora.irfft_f32)(X, self.n, nthreads)
            dt = time.perf_counter() - t0
        return dt, self.alg_bytes * nb / self.batch


WORKLOAD = [
    ("C2C f32 N=10007 x64", "c2c", 10007, 64, False),
    ("C2C f32 N=10935 x64", "c2c", 10935, 64, False),
    ("Z2Z f64 N=10007 x64", "c2c", 10007, 64, True),
    ("Z2Z f64 N=10935 x64", "c2c", 10935, 64, True),
    ("R2C->C2R f32 N=6000 x128", "real", 6000, 128, False),
    ("D2Z->Z2D f64 N=6000 x128", "real", 6000, 128, True),
]

SUITE = [  # the other configurations, reported with --suite (not the headline line)
    [("C2C f32 N=4096 x256", "c2c", 4096, 256, False)],
    [("C2C f32 N=2^20 x16", "c2c", 1 << 20, 16, False)],
    [("C2C f32 N=2^24 x1", "c2c", 1 << 24, 1, False)],
    [("C2C f32 N=1024 x65536", "c2c", 1024, 65536, False)],
    [("C2C f32 N=1024 x65536 FP16 hi+lo matrix cores", "c2c", 1024, 65536, False, 2)],
    [("C2C f32 N=1024 x65536 BF16 hi+lo matrix cores", "c2c", 1024, 65536, False, 1)],
    [("R2C->C2R f32 N=2^20 x16", "real", 1 << 20, 16, False)],
    [("2D R2C->C2R f32 4096x4096", "real", (4096, 4096), 1, False)],
    [("2D Z2Z f64 2048x2048", "c2c", (2048, 2048), 1, True)],
    [("3D C2C f32 256x256x256", "c2c", (256, 256, 256), 1, False)],
    [("3D R2C->C2R f32 256x256x256", "real", (256, 256, 256), 1, False)],
]


def kernel_symbol(label: str) -> str:
    """Step label -> the kernel symbol rocprof reports.

    'k_pass<fp32,16> ...'            -> 'k_pass<float, 16>'
    'k_leaf<fp32,16>[16x16x16] ...'   -> 'k_leaf<float, 16, 16, 16, 16>'
    'k_stage<fp32,16,8>[16x10] ...'   -> 'k_stage<float, 16, 8, 16, 10>'
    """
    head = label.split(" ")[0]
    fam, _, rest = head.partition("<")
    if not rest:
        return fam
    targs, _, rad = rest.partition(">")
    parts = targs.split(",")
    parts[0] = {"fp32": "float", "fp64": "double"}.get(parts[0], parts[0])
    if rad.startswith("["):
        parts += rad.strip("[]").split("x")
    return f"{fam}<{', '.join(parts)}>"


def hw_queues() -> int:
    """Hardware queues HIP gives one process (GPU_MAX_HW_QUEUES, default 4):
    streams beyond that share a queue and their kernels serialise."""
    try:
        return max(1, int(os.environ.get("GPU_MAX_HW_QUEUES", "4")))
    except ValueError:
        return 4


def assign_streams(cases, steps, torch, stream, nq):
    """Longest-first assignment of the transform families to nq streams by their
    device time alone (HIP events over `steps` steps on one stream)."""
    times = []
    for c in cases:
        c.stream = stream
        c.bind(stream.cuda_stream)
        with torch.cuda.stream(stream):
            c.run(0)
            e0 = torch.cuda.Event(enable_timing=True)
            e1 = torch.cuda.Event(enable_timing=True)
            e0.record(stream)
            for s in range(steps):
                c.run(s)
            e1.record(stream)
        e1.synchronize()
        times.append(e0.elapsed_time(e1) / steps)
    return lpt(times, nq), times


def lpt(times, nq):
    """Longest-first assignment of jobs with the given times to nq streams."""
    load = [0.0] * nq
    slot = [0] * len(times)
    for i in sorted(range(len(times)), key=lambda i: -times[i]):
        q = min(range(nq), key=lambda k: load[k])
        slot[i] = q
        load[q] += times[i]
    return slot


def pick_streams(cases, times, torch, dev, nq_max, steps=100):
    """Number of streams for the families: every count from 1 to nq_max with the
    longest-first assignment and two fresh stream sets per count, timed over
    `steps` steps (best of three regions); the fastest wins.  Returns
    (streams, slot, {count.set: us per step})."""
    table = {}
    best = None
    for nq in range(1, nq_max + 1):
        slot = lpt(times, nq)
        # two fresh stream sets per count: the runtime maps streams to its
        # hardware queues by itself, and two streams that share a queue run
        # their kernels one after the other
        for trial in range(1 if nq == 1 else 2):
            streams = [torch.cuda.Stream(device=dev) for _ in range(nq)]
            for i, c in enumerate(cases):
                c.stream = streams[slot[i]]
                c.bind(c.stream.cuda_stream)
            for s in range(10):
                for c in cases:
                    c.run(s)
            t = float("inf")
            for _ in range(3):
                torch.cuda.synchronize()
                t0 = time.perf_counter()
                for s in range(steps):
                    for c in cases:
                        c.run(s)
                torch.cuda.synchronize()
                t = min(t, time.perf_counter() - t0)
            table[f"{nq}" if nq == 1 else f"{nq}.{trial}"] = round(t / steps * 1e6, 1)
            if best is None or t < best[0]:
                best = (t, streams, slot)
    return best[1], best[2], table


def measure(cases, steps, warmup, torch, ff, stream, case_streams, slot):
    """Three timed regions of the same K steps:
    1. the transform families spread over the streams of case_streams (one per
       hardware queue, slot[i] = stream of case i), launches only (-> value);
    2. all families on one stream, launches only (-> single_stream_gbps);
    3. one stream with HIP events recorded around every launch (-> roofline)."""

    def bind(per_case):
        for i, c in enumerate(cases):
            st = case_streams[slot[i]] if per_case else stream
            c.stream = st
            c.bind(st.cuda_stream)

    def region(first_step):
        # each plan is bound to its stream (flagfftSetStream): no torch stream context needed
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for s in range(steps):
            for c in cases:
                c.run(first_step + s)
        torch.cuda.synchronize()
        return time.perf_counter() - t0

    bind(True)
    for s in range(max(warmup, 20)):
        for c in cases:
            c.run(s)
    t_multi = region(warmup)
    bind(False)
    for c in cases:
        c.run(0)
    t_single = region(warmup + steps)
    for c in cases:  # per-launch events from here on
        for p in c.plans:
            ff.check("flagfftxSetProfiling", ff.set_profiling(p.handle, True))
    region(warmup + 2 * steps)
    launches = {}
    per_case = {}
    for c in cases:
        case_ms = 0.0
        for pi, p in enumerate(c.plans):
            direction = ff.FORWARD if (c.kind == "c2c" or pi == 0) else ff.INVERSE
            for label, ms, calls, nbytes in ff.get_profile(p.handle, direction):
                key = (c.name, pi, label)
                launches[key] = (ms, calls, nbytes)
                case_ms += ms
        per_case[c.name] = {"alg_bytes": c.alg_bytes, "device_ms_per_exec": case_ms / steps,
                            "gbps": c.alg_bytes / (case_ms / steps * 1e-3) / 1e9 if case_ms > 0 else None}
    return t_multi, t_single, launches, per_case


TRAFFIC_FILE = os.path.join(ROOT, "profiles", "hbm_traffic_per_kernel.json")


def traffic_of(symbol: str):
    """HBM bytes per dispatch of `symbol` from the committed rocprofv3 counter summary, or None."""
    try:
        with open(TRAFFIC_FILE) as f:
            table = json.load(f)
    except (OSError, ValueError):
        return None
    row = table.get("kernels", {}).get(symbol)
    return None if row is None else int(row["hbm_bytes_per_dispatch"])


def counter_means(csv_path: str, counter: str):
    """Mean value of `counter` per kernel symbol in a rocprofv3 counter_collection.csv."""
    import csv

    sums, counts = {}, {}
    with open(csv_path) as f:
        for r in csv.DictReader(f):
            if r["Counter_Name"] != counter:
                continue
            name = r["Kernel_Name"].split("(")[0].replace("void ", "").replace("flagfft::dev::", "")
            sums[name] = sums.get(name, 0.0) + float(r["Counter_Value"])
            counts[name] = counts.get(name, 0) + 1
    return {k: sums[k] / counts[k] for k in sums}


def write_traffic(args):
    """Combine FETCH_SIZE / WRITE_SIZE runs (KiB per dispatch) of the bench and of a
    plain copy of known size into per-kernel HBM bytes per dispatch."""
    fetch = counter_means(args.fetch_csv, "FETCH_SIZE")
    write = counter_means(args.write_csv, "WRITE_SIZE")
    cf = counter_means(args.copy_fetch_csv, "FETCH_SIZE")
    cw = counter_means(args.copy_write_csv, "WRITE_SIZE")
    copy_name = max(cf, key=lambda k: cf[k])  # the calibration copy dominates its run
    half = args.copy_bytes / 2  # the copy reads and writes copy_bytes / 2 each
    f_factor = half / (cf[copy_name] * 1024.0)
    w_factor = half / (cw[copy_name] * 1024.0)
    kernels = {}
    for k in fetch:
        if k in write and k.startswith(("k_", "flagfft")):
            kernels[k] = {"fetch_kib": round(fetch[k], 1), "write_kib": round(write[k], 1),
                          "hbm_bytes_per_dispatch": int(fetch[k] * 1024 * f_factor + write[k] * 1024 * w_factor)}
    out = {"fetch_factor": round(f_factor, 4), "write_factor": round(w_factor, 4), "calibration_kernel": copy_name,
           "calibration_bytes": args.copy_bytes, "kernels": kernels}
    with open(args.write_traffic, "w") as f:
        json.dump(out, f, indent=1, sort_keys=True)
    print(json.dumps({"fetch_factor": out["fetch_factor"], "write_factor": out["write_factor"],
                      "kernels": len(kernels)}))


def roofline_of(launches):
    """Dominant kernel symbol by total device time; achieved = bytes / mean duration."""
    by_sym = {}
    for (case, pi, label), (ms, calls, nbytes) in launches.items():
        sym = kernel_symbol(label)
        d = by_sym.setdefault(sym, {"ms": 0.0, "calls": 0, "bytes": 0, "labels": set()})
        d["ms"] += ms
        d["calls"] += calls
        d["bytes"] += nbytes * calls
        d["labels"].add(f"{case}: {label}")
    sym, d = max(by_sym.items(), key=lambda kv: kv[1]["ms"])
    avg_ms = d["ms"] / d["calls"]
    bytes_per_launch = d["bytes"] / d["calls"]
    achieved = bytes_per_launch / (avg_ms * 1e-3) / 1e9
    traffic = traffic_of(sym)
    roof = {
        "bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
        "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": traffic,
        "kernel": sym, "avg_launch_us": round(avg_ms * 1e3, 3), "alg_bytes_per_launch": int(bytes_per_launch),
        "launches": d["calls"], "share_of_device_time": round(d["ms"] / sum(v["ms"] for v in by_sym.values()), 3),
    }
    if traffic:
        # HBM bytes the counters see per launch over the same mean duration: how
        # close the kernel's actual traffic comes to the HBM roof
        roof["traffic_gbps"] = round(traffic / (avg_ms * 1e-3) / 1e9, 1)
        roof["traffic_frac"] = round(roof["traffic_gbps"] / HBM_PEAK_GBS, 4)
    return roof, {k: {"ms": round(v["ms"], 3), "calls": v["calls"], "avg_us": round(v["ms"] / max(v["calls"], 1) * 1e3, 2)}
        for k, v in by_sym.items()}


def cpu_baseline(cases, budget_s=12.0):
    from oracle import oracle as ora

    nthreads = int(os.environ.get("OMP_NUM_THREADS", "0") or 0) or (os.cpu_count() or 1)
    nthreads = max(1, min(nthreads, os.cpu_count() or 1))
    frac = 0.25
    total_t = 0.0
    total_b = 0.0
    reps = 0
    t_start = time.perf_counter()
    while reps < 1 or time.perf_counter() - t_start < budget_s:
        for c in cases:
            dt, b = c.cpu_step(ora, nthreads, frac)
            total_t += dt
            total_b += b
        reps += 1
        if reps >= 200:
            break
    return {"value": round(total_b / total_t / 1e9, 3), "unit": "GB/s", "cores": nthreads, "kind": "port",
            "sample": f"{reps} x the same step with each batch cut to {int(frac * 100)}% "
                      f"(oracle fast FFT, FP32 for f32 cases, FP64 for f64 cases, {total_t:.1f} s wall)"}


def _is_prime(n: int) -> bool:
    if n < 2:
        return False
    d = 2
    while d * d <= n:
        if n % d == 0:
            return False
        d += 1
    return True


def serving_trace(nreq: int, seed: int, npool: int = 48):
    """Requests of the serving configuration: length log-uniform in [64, 2^20] with
    20% replaced by the nearest prime, batch uniform in [1, 512], type 70% C2C f32,
    20% R2C f32, 10% Z2Z f64.  Lengths come from a pool of `npool` such draws so
    that plans and tables are reused, as a serving system's would be."""
    rng = np.random.default_rng(seed)
    pool = []
    for _ in range(npool):
        n = int(round(float(np.exp(rng.uniform(np.log(64), np.log(1 << 20))))))
        if rng.uniform() < 0.2:
            while not _is_prime(n):
                n += 1
        pool.append(n)
    reqs = []
    for _ in range(nreq):
        u = rng.uniform()
        t = "C2C" if u < 0.7 else ("R2C" if u < 0.9 else "Z2Z")
        n = pool[int(rng.integers(len(pool)))]
        if t == "R2C" and n % 2:
            n += 1  # real transforms of the pool's odd lengths use the next even length
        reqs.append((t, n, int(rng.integers(1, 513))))
    return reqs


def serving_leg(torch, ff, dev, nreq: int, seed: int = 2024, budget_gib: float = 48.0):
    """Poisson stream through the server in BATCHED and PLAIN modes at the same offered rate."""
    import threading

    T = ff.Type
    reqs = serving_trace(nreq, seed)
    esz = {"C2C": 8, "R2C": 4, "Z2Z": 16}
    need = {t: max([n * b * esz[t] for tt, n, b in reqs if tt == t] or [0]) for t in esz}
    inputs = {t: torch.randn(max(need[t] // 4, 1), dtype=torch.float32, device=dev) for t in esz}
    torch.cuda.synchronize()
    tmap = {"C2C": T.C2C, "R2C": T.R2C, "Z2Z": T.Z2Z}
    out = {}

    def run(mode, rate):
        with ff.Server(int(budget_gib * (1 << 30)), max_inflight=3, mode=mode) as srv:
            for t, n, _ in set((t, n, 0) for t, n, b in reqs):
                srv.prepare(tmap[t], n, ff.FORWARD)
            tickets = []
            done = threading.Event()

            def reaper():
                i = 0
                while i < nreq:
                    if i < len(tickets):
                        srv.wait(tickets[i])
                        srv.release(tickets[i])
                        i += 1
                    else:
                        time.sleep(20e-6)
                done.set()

            th = threading.Thread(target=reaper, daemon=True)
            th.start()
            rng = np.random.default_rng(seed + 1)
            gaps = rng.exponential(1.0 / rate, nreq) if rate else np.zeros(nreq)
            t0 = time.perf_counter()
            t_next = t0
            for (t, n, b), g in zip(reqs, gaps):
                t_next += g
                while time.perf_counter() < t_next:
                    pass
                tickets.append(srv.submit(tmap[t], n, b, inputs[t].data_ptr(), ff.FORWARD))
            done.wait()
            th.join()
            return srv.stats()

    # capacity: every request submitted at once; the plain-launch capacity sets the offered rate
    caps = {name: run(mode, 0.0) for name, mode in (("batched", ff.SERVE_BATCHED), ("plain", ff.SERVE_PLAIN))}
    cap_rate = caps["plain"]["completed"] / caps["plain"]["span_s"]
    rate = 0.8 * cap_rate
    for name, mode in (("batched", ff.SERVE_BATCHED), ("plain", ff.SERVE_PLAIN)):
        s = run(mode, rate)
        out[name] = {"transforms_per_s": round(s["transforms"] / s["span_s"], 1),
                     "requests_per_s": round(s["completed"] / s["span_s"], 1),
                     "gbps": round(s["bytes"] / s["span_s"] / 1e9, 1),
                     "p50_ms": round(s["p50_ms"], 3), "p99_ms": round(s["p99_ms"], 3),
                     "queue_launches": s["queue_launches"], "queued_requests": s["queued_requests"]}
    out["capacity_transforms_per_s"] = {k: round(v["transforms"] / v["span_s"], 1) for k, v in caps.items()}
    out["capacity_gbps"] = {k: round(v["bytes"] / v["span_s"] / 1e9, 1) for k, v in caps.items()}
    out["offered_requests_per_s"] = round(rate, 1)
    out["requests"] = nreq
    out["plain_capacity_requests_per_s"] = round(cap_rate, 1)
    out["trace"] = ("capacity: all requests at once; latency: Poisson arrivals at 80% of the plain-launch "
                    "capacity; N log-uniform [64, 2^20] from a pool "
                    "of 48 lengths (20% prime); batch U[1,512]; 70% C2C f32, 20% R2C f32, 10% Z2Z f64")
    return out


def accuracy_sample(case, torch, rows: int = 64):
    """Relative L2 error of the first rows of one exec against the FP64 oracle."""
    from oracle import oracle as ora

    b = case.sets[0]
    case.run(0)
    torch.cuda.synchronize()
    if isinstance(case.n, tuple):  # 2D / 3D: the whole transform
        x = b[0].cpu().numpy().astype(np.complex128)
        y = b[1].cpu().numpy()
        ref = ora.fft2(x[None]) if len(case.n) == 2 else ora.fft3(x[None])
        return float(f"{ora.rel_l2(y.reshape(1, -1), ref.reshape(1, -1))[0]:.3e}")
    x = b[0][:rows].cpu().numpy().astype(np.complex128)
    y = b[1][:rows].cpu().numpy()
    return float(f"{ora.rel_l2(y, ora.fft(x))[0]:.3e}")


def run_workload(spec, args, torch, ff, dev, stream, seed):
    # one stream per transform family for the throughput region (default priority)
    cases = [Case(*row[:5], dev, ff, torch, *row[5:]) for row in spec]
    footprint = sum(c.footprint() for c in cases)
    nsets = max(1, -(-2 * MALL_BYTES // max(footprint, 1)) + 1)
    nsets = min(nsets, 64)
    for i, c in enumerate(cases):
        c.allocate(nsets, seed + i)
    torch.cuda.synchronize()
    nq = min(hw_queues(), len(cases))
    slot, alone_ms = assign_streams(cases, max(args.warmup, 3), torch, stream, nq)
    if nq > 1:
        case_streams, slot, stream_table = pick_streams(cases, alone_ms, torch, dev, nq)
    else:
        case_streams, stream_table = [torch.cuda.Stream(device=dev)], {}
    t, t1, launches, per_case = measure(cases, args.steps, args.warmup, torch, ff, stream, case_streams, slot)
    alg = sum(c.alg_bytes for c in cases)
    value = args.steps * alg / t / 1e9
    roof, kernels = roofline_of(launches)
    pass_bytes = sum(nb * calls for (ms, calls, nb) in launches.values())
    return cases, {"value": value, "ms_per_step": t / args.steps * 1e3, "alg_bytes_per_step": alg,
                   "single_stream_gbps": args.steps * alg / t1 / 1e9, "streams": len(case_streams),
                   "stream_calibration_us_per_step": stream_table,
                   "pass_gbps": pass_bytes / t / 1e9, "roofline": roof, "kernels": kernels,
                   "per_case": per_case, "buffer_sets": nsets, "footprint_mb": round(nsets * footprint / 2**20)}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--suite", action="store_true", help="also measure the other configurations (stderr)")
    ap.add_argument("--serve-requests", type=int, default=300, help="requests of the serving leg (0: skip)")
    ap.add_argument("--copy-probe", type=int, default=0, help="only run a device copy of this many bytes 20x "
                    "(counter calibration run under rocprofv3)")
    ap.add_argument("--write-traffic", help="write the per-kernel HBM traffic summary to this file and exit")
    ap.add_argument("--fetch-csv")
    ap.add_argument("--write-csv")
    ap.add_argument("--copy-fetch-csv")
    ap.add_argument("--copy-write-csv")
    ap.add_argument("--copy-bytes", type=int, default=0)
    args = ap.parse_args()
    if args.write_traffic:
        write_traffic(args)
        return
    if args.gpus != 1:
        print(f"bench.py measures one GPU only: --gpus {args.gpus} is not supported "
              "(the path shards by batch with no communication; see the design notes)", file=sys.stderr)
        sys.exit(2)

    import torch

    if args.copy_probe:
        src = torch.empty(args.copy_probe // 2, dtype=torch.uint8, device="cuda:0").fill_(1)
        dst = torch.empty_like(src)
        for _ in range(20):
            dst.copy_(src)
        torch.cuda.synchronize()
        return

    from flagfft_amd import flagfft as ff

    dev = torch.device("cuda:0")
    stream = torch.cuda.Stream(device=dev)
    cases, res = run_workload(WORKLOAD, args, torch, ff, dev, stream, 1234)

    cpu = None if args.no_cpu else cpu_baseline(cases)
    transforms_per_step = sum(c.batch * (1 if c.kind == "c2c" else 2) for c in cases)
    serving = None
    if args.serve_requests > 0:
        del cases
        torch.cuda.empty_cache()
        serving = serving_leg(torch, ff, dev, args.serve_requests)
        cases = []
    line = {
        "metric": METRIC,
        "value": round(res["value"], 2),
        "unit": "GB/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(res["ms_per_step"], 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32+f64",
        "data": "synthetic",
        "config": {
            "workload": "arbitrary-length 1D: C2C f32 + Z2Z f64 forward at N=10007 (prime) and N=10935=3^7*5, "
                        "batch 64, N(0,1); R2C->C2R f32 + D2Z->Z2D f64 at N=6000, batch 128, U[-1,1]",
            "transforms_per_step": transforms_per_step,
            "alg_bytes_per_step": res["alg_bytes_per_step"],
            "hbm_frac_of_peak": round(res["value"] / HBM_PEAK_GBS, 4),
            "pass_bytes_gbps": round(res["pass_gbps"], 2),
            "input_rotation": f"{res['buffer_sets']} buffer sets, {res['footprint_mb']} MiB",
            "parallelism": f"single GPU, transform families over {res['streams']} HIP streams "
                           f"(count calibrated over 1-{hw_queues()}), balanced by device time",
            "stream_calibration_us_per_step": res["stream_calibration_us_per_step"],
            "single_stream_gbps": round(res["single_stream_gbps"], 2),
        },
        "roofline": res["roofline"],
        "cpu_baseline": cpu,
        "serving": serving,
        "per_case": {k: {"gbps": round(v["gbps"], 1) if v["gbps"] else None,
                         "device_us": round(v["device_ms_per_exec"] * 1e3, 2)} for k, v in res["per_case"].items()},
        "kernels": res["kernels"],
    }
    if args.suite:
        for spec in SUITE:
            del cases
            torch.cuda.empty_cache()
            cases, r = run_workload(spec, args, torch, ff, dev, stream, 99)
            err = accuracy_sample(cases[0], torch) if cases[0].kind == "c2c" else None
            log(json.dumps({"suite": spec[0][0], "gbps": round(r["value"], 1), "rel_l2_vs_fp64": err,
                            "frac": round(r["value"] / HBM_PEAK_GBS, 3), "us_per_exec": round(r["ms_per_step"] * 1e3, 2),
                            "roofline": r["roofline"]}))
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
