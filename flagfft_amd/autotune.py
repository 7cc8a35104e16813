"""Autotuner that writes the tuned-plan database (table tuned_measurements).

For each (type, N, batch) it builds the default plan and a set of candidate
plans that differ in one of the planner's free choices (radix order of a
single-pass transform, super-radix split of a multi-pass chain, transforms
per workgroup, elements per thread), checks every candidate's output against
the default plan's on the same input (relative L2 within 10x the stated
accuracy bound), times each with HIP events on one stream (median of several
rounds of back-to-back executions), and appends the fastest to the database
through flagfftxTunedStore.  The default plan is always a candidate, so a
stored row is never slower than the untuned plan on the machine that wrote it.

The reference reads this table (src/utils/plan_json.cpp:170-207) but has no
writer (src/tune/tune.cpp:7-9 is a placeholder).

Usage (on a GPU):
  python -m flagfft_amd.autotune --case C2C:10935:64 --case R2C:6000:128 [--db PATH]
"""
from __future__ import annotations

import argparse
import itertools
import math
import os
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

from flagfft_amd import flagfft as ff  # noqa: E402

T = ff.Type
DOUBLE = {T.Z2Z, T.D2Z, T.Z2D}


def _factor(n: int):
    f, d = [], 2
    while d * d <= n:
        while n % d == 0:
            f.append(d)
            n //= d
        d += 1
    if n > 1:
        f.append(n)
    return f


def _radices(n: int):
    """First-fit-decreasing packing of the prime factors into radices <= 16,
    with 17 and 19 as radices of their own (the planner's default)."""
    bins = []
    for p in sorted(_factor(n), reverse=True):
        if p > 19:
            return None
        for i, b in enumerate(bins):
            if b * p <= 16:
                bins[i] = b * p
                break
        else:
            bins.append(p)
    return sorted(bins, reverse=True)


def _splits(n: int, cap: int, k: int):
    out = []

    def rec(rest, k, hi, cur):
        if k == 1:
            if 2 <= rest <= min(cap, hi) and _radices(rest):
                out.append(cur + [rest])
            return
        for d in range(min(cap, hi, rest), 1, -1):
            if rest % d == 0 and _radices(d):
                rec(rest // d, k - 1, d, cur + [d])

    rec(n, k, cap, [])
    return out


def candidates(type_, n: int, batch: int, describe: str):
    """Planner overrides worth timing for this transform (the default plan is {})."""
    cands = [{}]
    dbl = type_ in DOUBLE
    core_n = n // 2 if type_ in (T.R2C, T.D2Z, T.C2R, T.Z2D) and n % 2 == 0 else n
    if "lds_resident=1" in describe:
        # LDS Bluestein forms: 1-4 kernel variants (whole convolution per
        # workgroup), 5 whole convolution, 6 / 7 split over 2 / 4 workgroups
        for units in range(1, 8):
            cands.append({"units": units})
    elif "Leaf(" in describe:
        rad = _radices(core_n) or []
        orders = sorted(set(itertools.permutations(rad)), reverse=True)[:6]
        for o in orders:
            cands.append({"radices": list(o)})
        for units in (1, 2, 4, 8):
            cands.append({"units": units})
        for ept in ((8, 16) if dbl else (8, 16, 32)):
            cands.append({"ept": ept})
    if "MultiPassStockham(" in describe:
        cap = 2048 if dbl else 4096
        splits = _splits(core_n, cap, 2) + _splits(core_n, cap, 3)
        splits.sort(key=lambda s: (len(s), -min(s)))
        for s in splits[:8]:
            cands.append({"passes": s})
            cands.append({"passes": s[::-1]})
        # wider tiles of the specialised super-stages (longer contiguous runs;
        # 1000 NCH + C: C units in NCH exchange chunks)
        for units in (8, 16, 32, 64, 2016, 2032, 2064):
            cands.append({"units": units})
            for s in splits[:3]:
                cands.append({"passes": s, "units": units})
        cands.append({"stage_form": 2})  # LDS super-stages only
    uniq, seen = [], set()
    for c in cands:
        key = repr(sorted(c.items()))
        if key not in seen:
            seen.add(key)
            uniq.append(c)
    return uniq


class Runner:
    def __init__(self, torch, type_, n: int, batch: int, seed: int = 7):
        self.torch, self.type, self.n, self.batch = torch, type_, n, batch
        dev = torch.device("cuda:0")
        g = torch.Generator(device=dev)
        g.manual_seed(seed)
        dbl = type_ in DOUBLE
        cdt = torch.complex128 if dbl else torch.complex64
        rdt = torch.float64 if dbl else torch.float32
        h = n // 2 + 1
        if type_ in (T.C2C, T.Z2Z):
            shapes = ((batch, n, cdt), (batch, n, cdt))
        elif type_ in (T.R2C, T.D2Z):
            shapes = ((batch, n, rdt), (batch, h, cdt))
        else:
            shapes = ((batch, h, cdt), (batch, n, rdt))
        b, l0, dt0 = shapes[0]
        b, l1, dt1 = shapes[1]
        esz = {torch.complex128: 16, torch.complex64: 8, torch.float64: 8, torch.float32: 4}
        self.bytes = b * (l0 * esz[dt0] + l1 * esz[dt1])
        # input / output sets rotated between executions, together larger than
        # twice the 256 MB Infinity Cache, so that each execution reads HBM as
        # in a stream of distinct requests
        nsets = max(1, min(8, -(-(640 << 20) // self.bytes)))
        self.xs = [torch.randn(b, l0, dtype=dt0, device=dev, generator=g) for _ in range(nsets)]
        self.ys = [torch.empty(b, l1, dtype=dt1, device=dev) for _ in range(nsets)]
        self.stream = torch.cuda.Stream(device=dev)

    def exec(self, p, k: int = 0):
        t = self.type
        i, o = self.xs[k % len(self.xs)].data_ptr(), self.ys[k % len(self.ys)].data_ptr()
        if t == T.C2C:
            p.c2c(i, o, ff.FORWARD)
        elif t == T.Z2Z:
            p.z2z(i, o, ff.FORWARD)
        elif t == T.R2C:
            p.r2c(i, o)
        elif t == T.D2Z:
            p.d2z(i, o)
        elif t == T.C2R:
            p.c2r(i, o)
        else:
            p.z2d(i, o)

    def measure(self, p, rounds: int = 7, per_round: int = 20):
        torch = self.torch
        p.set_stream(self.stream.cuda_stream)
        with torch.cuda.stream(self.stream):
            for k in range(3):
                self.exec(p, k)
            self.exec(p, 0)
            self.stream.synchronize()
            out = self.ys[0].clone()
            times = []
            for _ in range(rounds):
                e0 = torch.cuda.Event(enable_timing=True)
                e1 = torch.cuda.Event(enable_timing=True)
                e0.record(self.stream)
                for k in range(per_round):
                    self.exec(p, k)
                e1.record(self.stream)
                e1.synchronize()
                times.append(e0.elapsed_time(e1) * 1e3 / per_round)
        return float(np.median(times)), out


def tune_case(torch, type_, n: int, batch: int, store: bool = True, log=print):
    ff.set_tuning()
    os.environ["FLAGFFT_TUNE_DISABLE"] = "1"  # measure the untuned planner's candidates
    try:
        run = Runner(torch, type_, n, batch)
        r, text = ff.describe_host(1, [n], type_, batch)
        ff.check("flagfftxPlanDescribe", r)
        bound = (1e-14 if type_ in DOUBLE else 2e-6) * max(math.log2(n), 1.0) * 10
        results = []
        ref = None
        for cand in candidates(type_, n, batch, text):
            ff.set_tuning(**cand) if cand else ff.set_tuning()
            try:
                with ff.Plan.one_d(n, type_, batch) as p:
                    us, out = run.measure(p)
            except ff.FlagFFTError as e:
                log(f"  skip {cand}: {e}")
                continue
            finally:
                ff.set_tuning()
            if ref is None:
                ref = out
                err = 0.0
            else:
                err = float(torch.linalg.vector_norm(out - ref) / torch.linalg.vector_norm(ref))
            ok = err <= bound
            log(f"  {type_.name} n={n} batch={batch} {cand or 'default'}: {us:.2f} us rel={err:.2e}"
                f"{'' if ok else ' REJECTED'}")
            if ok:
                results.append((us, cand))
        default_us = results[0][0] if results and results[0][1] == {} else float("nan")
        best_us, best = min(results, key=lambda r: r[0])
        gbps = run.bytes / (best_us * 1e-6) / 1e9
        if store:
            os.environ.pop("FLAGFFT_TUNE_DISABLE", None)
            ff.check("flagfftxTunedStore", ff.tuned_store(n, type_, batch, best_us, gbps, default_us, **best))
        return {"type": type_.name, "n": n, "batch": batch, "best": best, "best_us": best_us,
                "default_us": default_us, "gbps": gbps}
    finally:
        os.environ.pop("FLAGFFT_TUNE_DISABLE", None)


def main():
    ap = argparse.ArgumentParser(description=__doc__.splitlines()[0])
    ap.add_argument("--case", action="append", required=True, help="TYPE:N:BATCH, e.g. C2C:10935:64")
    ap.add_argument("--db", help="database file (default: $FLAGFFT_TUNE_DB or beside libflagfft.so)")
    ap.add_argument("--dry-run", action="store_true", help="time candidates without storing")
    args = ap.parse_args()
    if args.db:
        os.environ["FLAGFFT_TUNE_DB"] = args.db
    import torch

    for spec in args.case:
        tname, n, batch = spec.split(":")
        res = tune_case(torch, T[tname], int(n), int(batch), store=not args.dry_run)
        print(f"best {res['type']} n={res['n']} batch={res['batch']}: {res['best'] or 'default'} "
              f"{res['best_us']:.2f} us (default {res['default_us']:.2f} us), {res['gbps']:.0f} GB/s", flush=True)
    if not args.dry_run:
        print("database:", ff.tuned_db_path())


if __name__ == "__main__":
    main()
