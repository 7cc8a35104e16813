"""Benchmark of the FFT execution path behind include/flagfft.h on one MI355X.

Headline workload (the large C2C/R2C configuration; one "step" = one pass of
the hot path over one batch of synthetic input, every transform through the
C ABI of libflagfft.so):
  C2C f32 forward N=2^24 x 1 and N=2^20 x 16 (re, im iid N(0,1));
  2D R2C -> C2R f32 4096 x 4096 (real inputs iid N(0,1));
  2D Z2Z f64 forward 2048 x 2048 (re, im iid N(0,1)).
Inputs rotate over buffer sets larger than 2x the 256 MB Infinity Cache so the
timed region reads HBM.

value          algorithmic HBM GB/s of the whole step: sum over transforms of
               input bytes + output bytes, divided by wall time (device
               synchronised on both sides of the timed region); the transform
               families are spread longest-first by their device time over 1
               to 4 HIP streams (4 hardware queues per process), the count
               chosen by a short calibration region before the timed one
config.pass_bytes_gbps
               the same region counted as passes x bytes (every launch's read
               + write, workspaces included): the north-star accounting
roofline       dominant kernel (by total device time): algorithmic bytes per
               launch / mean launch duration from HIP events recorded on the
               plan's stream around every launch of a second timed region of
               the same K steps; traffic = counter-measured HBM bytes per
               launch (committed rocprofv3 summary)
cpu_baseline   the oracle's fast FFT (oracle/oracle.c) on the host cores over
               the same step (median of 10 repetitions)
secondary      the other configurations under their own keys: the
               arbitrary-length 1D mix, C2C f32 4096 x 256 forward + inverse,
               the precision form (1024 x 65536 vector / FP16 hi+lo / BF16
               hi+lo with their errors against the FP64 oracle) and the
               serving leg (Poisson stream, batched vs plain launches)

Usage: python bench.py [--gpus 1] [--steps K] [--warmup W] [--no-cpu]
                       [--headline-only] [--serve-requests R] [--suite]
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
    """One transform family of a workload: plans, rotating device buffers, one exec.

    kind: "c2c" (forward), "c2c_rt" (forward then inverse), "real" (R2C then C2R).
    n: an int (1D, `batch` rows) or a tuple (2D / 3D row-major, one transform)."""

    def __init__(self, name, kind, n, batch, dbl, dev, ff, torch, compute=0, real_dist=None):
        self.name, self.kind, self.n, self.batch, self.dbl = name, kind, n, batch, dbl
        self.compute = compute
        self.ff, self.torch, self.dev = ff, torch, dev
        # real inputs: iid N(0,1) for the 2D / 3D transforms, U[-1,1] for the 1D ones
        self.real_dist = real_dist or ("normal" if isinstance(n, tuple) else "uniform")
        cdt = torch.complex128 if dbl else torch.complex64
        rdt = torch.float64 if dbl else torch.float32
        esz = 16 if dbl else 8
        T = ff.Type
        F, I = ff.FORWARD, ff.INVERSE
        if isinstance(n, tuple):  # 2D / 3D (row-major, last axis contiguous), one transform per exec
            lead, nl = n[:-1], n[-1]
            h = nl // 2 + 1
            outer = int(np.prod(lead))
            self.batch = 1

            def plan(t):
                return ff.Plan.two_d(n[0], n[1], t) if len(n) == 2 else ff.Plan.three_d(n[0], n[1], n[2], t, 1)

            if kind == "c2c":
                self.plans = [plan(T.Z2Z if dbl else T.C2C)]
                self.profile = [(0, F)]
                self.alg_bytes = outer * nl * esz * 2
                self.shapes = [(n, cdt), (n, cdt)]
            else:
                self.plans = [plan(T.D2Z if dbl else T.R2C), plan(T.Z2D if dbl else T.C2R)]
                self.profile = [(0, F), (1, I)]
                self.alg_bytes = 2 * (outer * nl * esz // 2 + outer * h * esz)
                self.shapes = [(n, rdt), (lead + (h,), cdt), (n, rdt)]
            self.sets = []
            return
        h = n // 2 + 1
        if kind in ("c2c", "c2c_rt"):
            self.plans = [ff.Plan.one_d_compute(n, T.Z2Z if dbl else T.C2C, batch, compute)]
            self.profile = [(0, F)] if kind == "c2c" else [(0, F), (0, I)]
            self.alg_bytes = batch * n * esz * 2 * (1 if kind == "c2c" else 2)
            self.shapes = [((batch, n), cdt), ((batch, n), cdt)] + ([((batch, n), cdt)] if kind == "c2c_rt" else [])
        else:  # real roundtrip R2C -> C2R
            self.plans = [ff.Plan.one_d(n, T.D2Z if dbl else T.R2C, batch),
                          ff.Plan.one_d(n, T.Z2D if dbl else T.C2R, batch)]
            self.profile = [(0, F), (1, I)]
            self.alg_bytes = 2 * batch * (n * esz // 2 + h * esz)
            self.shapes = [((batch, n), rdt), ((batch, h), cdt), ((batch, n), rdt)]
        self.sets = []

    def transforms(self):
        return self.batch * (1 if self.kind == "c2c" else 2)

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
                    elif self.real_dist == "normal":
                        t = torch.randn(shape, dtype=dt, device=self.dev, generator=g)  # N(0, 1)
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
        if self.kind in ("c2c", "c2c_rt"):
            ex = p[0].z2z if self.dbl else p[0].c2c
            ex(b[0].data_ptr(), b[1].data_ptr(), self.ff.FORWARD)
            if self.kind == "c2c_rt":
                ex(b[1].data_ptr(), b[2].data_ptr(), self.ff.INVERSE)
        else:
            (p[0].d2z if self.dbl else p[0].r2c)(b[0].data_ptr(), b[1].data_ptr())
            (p[1].z2d if self.dbl else p[1].c2r)(b[1].data_ptr(), b[2].data_ptr())

    # ---- the same step on the host cores (the oracle's fast FFT) -------------
    def cpu_prepare(self, nthreads):
        """Host input of the whole step and the tables a CPU plan would build once."""
        rng = np.random.default_rng(abs(hash(self.name)) % (1 << 32))
        shape = self.n if isinstance(self.n, tuple) else (self.batch, self.n)
        if self.kind == "real":
            x = rng.standard_normal(shape) if self.real_dist == "normal" else rng.uniform(-1, 1, shape)
            self.cpu_x = x.astype(np.float64 if self.dbl else np.float32)
        else:
            x = rng.standard_normal(shape) + 1j * rng.standard_normal(shape)
            self.cpu_x = x.astype(np.complex128 if self.dbl else np.complex64)
        self.cpu_tw = None
        if not isinstance(self.n, tuple) and self.batch < nthreads and self.n >= (1 << 16) and self.n & (self.n - 1) == 0:
            # batch below the thread count: four-step over sqrt(N) x sqrt(N) so every core works
            lg = self.n.bit_length() - 1
            n1, n2 = 1 << (lg // 2), 1 << (lg - lg // 2)
            b = np.arange(n2)[:, None]
            k = np.arange(n1)[None, :]
            tw = np.exp(-2j * np.pi * ((b * k) % self.n) / self.n)
            self.cpu_tw = (n1, n2, tw.astype(np.complex128 if self.dbl else np.complex64))

    def _c2c_rows(self, ora, x, sign, nthreads):
        return (ora.fft if self.dbl else ora.fft_f32)(x, sign, nthreads)

    def _c2c_1d(self, ora, x, sign, nthreads):
        if self.cpu_tw is None:
            return self._c2c_rows(ora, x, sign, nthreads)
        n1, n2, tw = self.cpu_tw
        out = np.empty_like(x)
        for r in range(x.shape[0]):  # X[k1 + n1 k2] = sum_b w_N^(b k1) w_n2^(b k2) sum_a x[n2 a + b] w_n1^(a k1)
            B = np.ascontiguousarray(x[r].reshape(n1, n2).T)
            C = self._c2c_rows(ora, B, sign, nthreads)
            C *= tw if sign < 0 else np.conj(tw)
            E = self._c2c_rows(ora, np.ascontiguousarray(C.T), sign, nthreads)
            out[r] = np.ascontiguousarray(E.T).reshape(-1)
        return out

    def cpu_run(self, ora, nthreads):
        x = self.cpu_x
        if isinstance(self.n, tuple):
            if len(self.n) != 2:
                raise NotImplementedError("CPU baseline of rank-3 transforms")
            if self.kind == "c2c":
                if self.dbl:
                    return ora.fft2(x, -1, nthreads)
                y = self._c2c_rows(ora, x, -1, nthreads)
                return np.ascontiguousarray(self._c2c_rows(ora, np.ascontiguousarray(y.T), -1, nthreads).T)
            rf, irf = (ora.rfft, ora.irfft) if self.dbl else (ora.rfft_f32, ora.irfft_f32)
            X = rf(x, nthreads)  # rows, then columns over a transpose
            X = self._c2c_rows(ora, np.ascontiguousarray(X.T), -1, nthreads)
            X = self._c2c_rows(ora, X, 1, nthreads)  # inverse: columns, then rows
            return irf(np.ascontiguousarray(X.T), self.n[-1], nthreads)
        if self.kind == "real":
            rf, irf = (ora.rfft, ora.irfft) if self.dbl else (ora.rfft_f32, ora.irfft_f32)
            return irf(rf(x, nthreads), self.n, nthreads)
        y = self._c2c_1d(ora, x, -1, nthreads)
        if self.kind == "c2c_rt":
            y = self._c2c_1d(ora, y, 1, nthreads)
        return y


WORKLOAD = [  # headline: C2C/R2C at N = 2^20 - 2^24 and the 2D transforms
    ("C2C f32 N=2^24 x1", "c2c", 1 << 24, 1, False),
    ("C2C f32 N=2^20 x16", "c2c", 1 << 20, 16, False),
    ("2D R2C->C2R f32 4096x4096", "real", (4096, 4096), 1, False),
    ("2D Z2Z f64 2048x2048", "c2c", (2048, 2048), 1, True),
]
WORKLOAD_NAME = ("large C2C/R2C: C2C f32 forward N=2^24 x1 and N=2^20 x16, 2D R2C->C2R f32 4096x4096, "
                 "2D Z2Z f64 forward 2048x2048; N(0,1) inputs")

ARBITRARY = [  # arbitrary-length 1D (Rader/Bluestein primes, mixed radix, real roundtrips)
    ("C2C f32 N=10007 x64", "c2c", 10007, 64, False),
    ("C2C f32 N=10935 x64", "c2c", 10935, 64, False),
    ("Z2Z f64 N=10007 x64", "c2c", 10007, 64, True),
    ("Z2Z f64 N=10935 x64", "c2c", 10935, 64, True),
    ("R2C->C2R f32 N=6000 x128", "real", 6000, 128, False),
    ("D2Z->Z2D f64 N=6000 x128", "real", 6000, 128, True),
]

SMALL = [("C2C f32 N=4096 x256 forward+inverse", "c2c_rt", 4096, 256, False)]

PRECISION = [  # (name, compute mode)
    ("vector FP32", 0), ("FP16 hi+lo matrix cores", 2), ("BF16 hi+lo matrix cores", 1),
]

SUITE = [  # further configurations, reported with --suite (stderr)
    [("C2C f32 N=1024 x65536", "c2c", 1024, 65536, False)],
    [("R2C->C2R f32 N=2^20 x16", "real", 1 << 20, 16, False)],
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
        for pi, direction in c.profile:
            for label, ms, calls, nbytes in ff.get_profile(c.plans[pi].handle, direction):
                key = (c.name, pi, direction, label)
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
    for (case, pi, direction, label), (ms, calls, nbytes) in launches.items():
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
    # "bound" is the roof the kernel's arithmetic intensity puts it under (every
    # FFT kernel here is below the FP32 vector ridge); "limiter" says whether
    # the measured rate actually reaches that roof
    seen = roof.get("traffic_frac", roof["frac"])
    roof["limiter"] = "hbm" if seen >= 0.6 else "not hbm: issue or latency bound below 60% of the HBM roof"
    return roof, {k: {"ms": round(v["ms"], 3), "calls": v["calls"], "avg_us": round(v["ms"] / max(v["calls"], 1) * 1e3, 2)}
        for k, v in by_sym.items()}


def cpu_model() -> str:
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    import platform

    return platform.processor() or "unknown"


def cpu_baseline(cases, reps: int = 10):
    """The same step through the oracle's fast FFT on the host cores (FP32 for
    the f32 cases, FP64 for the f64 ones): one untimed repetition, then the
    median of `reps` timed ones."""
    from oracle import oracle as ora

    nthreads = int(os.environ.get("OMP_NUM_THREADS", "0") or 0) or (os.cpu_count() or 1)
    nthreads = max(1, min(nthreads, os.cpu_count() or 1))
    for c in cases:
        c.cpu_prepare(nthreads)
    alg = sum(c.alg_bytes for c in cases)
    times = []
    for r in range(reps + 1):
        t0 = time.perf_counter()
        for c in cases:
            c.cpu_run(ora, nthreads)
        if r > 0:
            times.append(time.perf_counter() - t0)
    med = float(np.median(times))
    return {"value": round(alg / med / 1e9, 3), "unit": "GB/s", "cores": nthreads, "kind": "port",
            "cpu": cpu_model(),
            "sample": f"the whole step ({alg / 2**20:.0f} MiB of algorithmic bytes) on {nthreads} host threads, "
                      f"median of {reps} repetitions after one untimed ({med * 1e3:.0f} ms per step, "
                      f"{sum(times):.1f} s timed); oracle fast FFT, FP32 for f32 cases, FP64 for f64 cases; "
                      "batch-1 2^24 as a threaded four-step, 2D as rows then columns over a transpose",
            "ms_per_step": round(med * 1e3, 1)}


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

    def run(mode, rate):
        with ff.Server(int(budget_gib * (1 << 30)), max_inflight=3, mode=mode) as srv:
            for t, n, _ in set((t, n, 0) for t, n, b in reqs):
                srv.prepare(tmap[t], n, ff.FORWARD)
            tickets = []
            done = threading.Event()

            def reaper():
                # a client that takes each result as it completes: the oldest
                # outstanding requests are polled every 100 us and released
                # when done, in whatever order they finish
                outstanding, seen, released = [], 0, 0
                while released < nreq:
                    while seen < len(tickets):
                        outstanding.append(tickets[seen])
                        seen += 1
                    keep = []
                    for i, tk in enumerate(outstanding):
                        if i < 64 and srv.poll(tk)[0]:
                            srv.release(tk)
                            released += 1
                        else:
                            keep.append(tk)
                    outstanding = keep
                    time.sleep(100e-6)
                done.set()

            th = threading.Thread(target=reaper, daemon=True)
            th.start()
            rng = np.random.default_rng(seed + 1)
            gaps = rng.exponential(1.0 / rate, nreq) if rate else np.zeros(nreq)
            t0 = time.perf_counter()
            t_next = t0
            for (t, n, b), g in zip(reqs, gaps):
                t_next += g
                # sleep (which lets the reaper thread run) until shortly
                # before the arrival time, then spin to it
                ahead = t_next - time.perf_counter() - 300e-6
                if ahead > 0:
                    time.sleep(ahead)
                while time.perf_counter() < t_next:
                    pass
                tickets.append(srv.submit(tmap[t], n, b, inputs[t].data_ptr(), ff.FORWARD))
            done.wait()
            th.join()
            return srv.stats()

    # the submitting thread and the reaper share the interpreter lock: switch
    # between them every 100 us (default 5 ms) so that neither holds the other up
    switch = sys.getswitchinterval()
    sys.setswitchinterval(100e-6)
    try:
        return _serving_runs(run, ff, nreq)
    finally:
        sys.setswitchinterval(switch)


def _serving_runs(run, ff, nreq):
    """Capacity and Poisson-latency runs of both server modes."""
    out = {}
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


def sample_rows(batch: int, per_part: int = 16):
    """Rows from the start, the middle and the end of a batch (every part of a
    kernel's grid and of a persistent kernel's loop)."""
    rows = set()
    for start in (0, batch // 4, batch // 2, batch - per_part):
        rows.update(range(max(0, start), min(batch, max(0, start) + per_part)))
    return np.array(sorted(rows))


def accuracy(case, torch):
    """Relative L2 errors of one exec of `case` (buffer set 0) against the FP64
    oracle: the forward output, and for round trips the second transform's
    output against n x the input (the inverse is unnormalised)."""
    from oracle import oracle as ora

    b = case.sets[0]
    case.run(0)
    torch.cuda.synchronize()
    out = {}

    def err(y, ref):
        return float(f"{ora.rel_l2(y, ref, y.shape[-1])[0]:.3e}")

    if isinstance(case.n, tuple):
        if len(case.n) != 2:
            return out
        x = b[0].cpu().numpy()
        npts = case.n[0] * case.n[1]
        if case.kind == "c2c":
            out["forward"] = err(b[1].cpu().numpy().reshape(1, -1), ora.fft2(x.astype(np.complex128)).reshape(1, -1))
        else:
            out["forward"] = err(b[1].cpu().numpy().reshape(1, -1), ora.rfft2(x.astype(np.float64)).reshape(1, -1))
            out["roundtrip"] = err(b[2].cpu().numpy().reshape(1, -1), npts * x.astype(np.float64).reshape(1, -1))
        return out
    rows = sample_rows(case.batch)
    rt = torch.as_tensor(rows, device=case.dev)
    x = b[0].index_select(0, rt).cpu().numpy()
    if case.kind == "real":
        out["forward"] = err(b[1].index_select(0, rt).cpu().numpy(), ora.rfft(x.astype(np.float64)))
        out["roundtrip"] = err(b[2].index_select(0, rt).cpu().numpy(), case.n * x.astype(np.float64))
    else:
        out["forward"] = err(b[1].index_select(0, rt).cpu().numpy(), ora.fft(x.astype(np.complex128)))
        if case.kind == "c2c_rt":
            out["roundtrip"] = err(b[2].index_select(0, rt).cpu().numpy(), case.n * x.astype(np.complex128))
    out["rows_checked"] = int(len(rows))
    return out


def run_workload(spec, args, torch, ff, dev, stream, seed):
    # one stream per transform family for the throughput region (default priority)
    cases = [Case(*row[:5], dev, ff, torch, *row[5:]) for row in spec]
    if args.graph:
        for c in cases:
            for p in c.plans:
                p.set_graph_mode(True)
    footprint = sum(c.footprint() for c in cases)
    nsets = max(1, -(-2 * MALL_BYTES // max(footprint, 1)) + 1)
    nsets = min(nsets, 64)
    for i, c in enumerate(cases):
        c.allocate(nsets, seed + i)
    torch.cuda.synchronize()
    nq = min(args.streams if args.streams > 0 else hw_queues(), len(cases))
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


def leg_summary(res):
    """The figures of one measured workload other than the headline."""
    return {"gbps": round(res["value"], 1), "frac_of_peak": round(res["value"] / HBM_PEAK_GBS, 4),
            "pass_bytes_gbps": round(res["pass_gbps"], 1), "us_per_step": round(res["ms_per_step"] * 1e3, 2),
            "streams": res["streams"], "roofline": res["roofline"],
            "per_case": per_case_summary(res)}


def per_case_summary(res):
    return {k: {"gbps": round(v["gbps"], 1) if v["gbps"] else None,
                "device_us": round(v["device_ms_per_exec"] * 1e3, 2)} for k, v in res["per_case"].items()}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-accuracy", action="store_true", help="skip the FP64-oracle error reports")
    ap.add_argument("--headline-only", action="store_true", help="measure the headline workload only")
    ap.add_argument("--graph", action="store_true", help="run every plan's launches as one hipGraph (flagfftxSetGraphMode)")
    ap.add_argument("--suite", action="store_true", help="also measure further configurations (stderr)")
    ap.add_argument("--serve-requests", type=int, default=300, help="requests of the serving leg (0: skip)")
    ap.add_argument("--streams", type=int, default=0,
                    help="streams for the families (0: calibrate over 1 to the hardware-queue count; 1: one stream, "
                         "every kernel alone on the GPU, as in the per-launch profiling runs)")
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

    def release(cases):
        for c in cases:
            for p in c.plans:
                p.destroy()
            c.sets.clear()
        torch.cuda.synchronize()
        torch.cuda.empty_cache()

    log("headline workload: " + WORKLOAD_NAME)
    cases, res = run_workload(WORKLOAD, args, torch, ff, dev, stream, 1234)
    errors = {} if args.no_accuracy else {c.name: accuracy(c, torch) for c in cases}
    log("headline measured; accuracy " + json.dumps(errors))
    cpu = None if args.no_cpu else cpu_baseline(cases)
    transforms_per_step = sum(c.transforms() for c in cases)
    release(cases)

    secondary = {}
    if not args.headline_only:
        for key, spec, seed in (("arbitrary_length_1d", ARBITRARY, 2001), ("c2c_f32_4096x256_roundtrip", SMALL, 2002)):
            log(f"secondary workload: {key}")
            cases, r = run_workload(spec, args, torch, ff, dev, stream, seed)
            secondary[key] = leg_summary(r)
            if not args.no_accuracy:
                secondary[key]["rel_l2_vs_fp64"] = {c.name: accuracy(c, torch) for c in cases}
            release(cases)
        prec = {}
        for name, mode in PRECISION:
            log(f"precision form: {name}")
            cases, r = run_workload([(f"C2C f32 N=1024 x65536 {name}", "c2c", 1024, 65536, False, mode)], args, torch,
                                    ff, dev, stream, 2003)
            prec[name] = {"gbps": round(r["value"], 1), "frac_of_peak": round(r["value"] / HBM_PEAK_GBS, 4),
                          "us_per_exec": round(r["ms_per_step"] * 1e3, 2),
                          "kernel": r["roofline"]["kernel"], "kernel_us": r["roofline"]["avg_launch_us"]}
            if not args.no_accuracy:
                prec[name]["rel_l2_vs_fp64"] = accuracy(cases[0], torch)
            release(cases)
        secondary["precision_form_1024x65536"] = prec
        if args.serve_requests > 0:
            log("serving leg")
            secondary["serving"] = serving_leg(torch, ff, dev, args.serve_requests)

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
            "workload": WORKLOAD_NAME,
            "transforms_per_step": transforms_per_step,
            "alg_bytes_per_step": res["alg_bytes_per_step"],
            "hbm_frac_of_peak": round(res["value"] / HBM_PEAK_GBS, 4),
            "pass_bytes_gbps": round(res["pass_gbps"], 2),
            "pass_bytes_frac_of_peak": round(res["pass_gbps"] / HBM_PEAK_GBS, 4),
            "input_rotation": f"{res['buffer_sets']} buffer sets, {res['footprint_mb']} MiB",
            "parallelism": f"single GPU, transform families over {res['streams']} HIP streams "
                           f"(count calibrated over 1-{hw_queues()}), balanced by device time",
            "stream_calibration_us_per_step": res["stream_calibration_us_per_step"],
            "single_stream_gbps": round(res["single_stream_gbps"], 2),
        },
        "roofline": res["roofline"],
        "cpu_baseline": cpu,
        "rel_l2_vs_fp64": errors,
        "per_case": per_case_summary(res),
        "kernels": res["kernels"],
        "secondary": secondary,
    }
    if args.suite:
        for spec in SUITE:
            cases, r = run_workload(spec, args, torch, ff, dev, stream, 99)
            err = None if args.no_accuracy else accuracy(cases[0], torch)
            log(json.dumps({"suite": spec[0][0], "gbps": round(r["value"], 1), "rel_l2_vs_fp64": err,
                            "frac": round(r["value"] / HBM_PEAK_GBS, 3), "us_per_exec": round(r["ms_per_step"] * 1e3, 2),
                            "pass_bytes_gbps": round(r["pass_gbps"], 1), "roofline": r["roofline"]}))
            release(cases)
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
