"""flagfft-cli for this library: `python -m flagfft_amd.cli bench ...` and `... tune ...`.

bench follows the reference's command line and report:
  options            src/cli_tools/flagfft-cli/main.cpp:27-43, validation :100-124
  buffers and input  src/cli_tools/bench/runner.cpp:25-80 (one allocation of
                     max(in, out) bytes per library, sin((i + 1) 0.173) fill,
                     Im of the DC / Nyquist bins zeroed for C2R / Z2D)
  timing             runner.cpp:229-303: warmup of both libraries, then
                     iterations alternating which one goes first, every exec
                     bracketed by events on one stream; median and p90
  report             src/cli_tools/bench/report.cpp:12-67 (table, and JSON
                     {"status", "command", "cases": [...]})
  exit codes         src/cli_tools/common/cli_utils.hpp:14-17: 0 passed,
                     1 failed (bad arguments), 2 runtime error, 77 skipped
                     (no GPU)
The comparator ("ref") is torch.fft, which runs hipFFT / rocFFT on ROCm, where
the reference runs cuFFT; --no-ref leaves it out.  Beyond the reference's
fields each JSON case carries a "bytes" block: the algorithmic bytes at the API
boundary (batch x (input + output)), the plan's launches and the bytes its
passes move, and the effective GB/s of both libraries; --check adds the
relative L2 distance between the two outputs.

Differences on purpose: real transforms of rank 2 are accepted (batch 1, the
reference's CLI refuses them); rank 3 runs through flagfftxPlan3d (any type,
batched; the reference's library answers FLAGFFT_NOT_SUPPORTED there); 1D
real in-place cases use the padded layout of 2(N/2+1) reals per row through
PlanMany, the in-place layout the library accepts
(src/exec/c_api_internal.cpp:74-89).

tune runs the autotuner (flagfft_amd.autotune) with the remaining arguments;
the reference's tune command is a placeholder (src/cli_tools/tune/tune.cpp:7-9).
"""
from __future__ import annotations

import json
import math
import sys

EXIT_PASSED, EXIT_FAILED, EXIT_RUNTIME_ERROR, EXIT_SKIPPED = 0, 1, 2, 77

APIS = ("c2c", "z2z", "r2c", "d2z", "c2r", "z2d")
COMPLEX = ("c2c", "z2z")
DOUBLE = ("z2z", "d2z", "z2d")
REAL_FWD = ("r2c", "d2z")
REAL_INV = ("c2r", "z2d")

USAGE = """Usage: python -m flagfft_amd.cli bench --rank 1|2|3 --shape <format> [options]
       python -m flagfft_amd.cli tune --case TYPE:N:BATCH [...]

Bench options:
  --rank 1|2|3             Transform rank (required)
  --shape A,B,...          Comma-separated shapes, x for dims
                           rank 1: --shape 16,32,64
                           rank 2: --shape 23x42,128x64
  --api c2c|z2z|r2c|d2z|c2r|z2d   FFT type (default: c2c)
  --batch N                Batch size for rank 1, complex rank 2 and rank 3 (default: 1)
  --direction forward|inverse       (default: forward)
  --placement out-of-place|in-place (default: out-of-place)
  --warmup N               Warmup iterations (default: 10)
  --iters N                Timed iterations (default: 100)
  --json                   Output JSON report
  --print-path             Include plan description in output
  --no-ref                 Time this library only (no torch.fft comparator)
  --check                  Report the relative L2 distance to the comparator's output
"""


class CliError(Exception):
    def __init__(self, message: str, code: int):
        super().__init__(message)
        self.code = code


def fail(message: str):
    raise CliError(message, EXIT_FAILED)


def parse_positive(name: str, value: str, allow_zero: bool = False) -> int:
    try:
        v = int(value)
    except ValueError:
        fail(f"invalid value for {name}")
    if v < 0 or (v == 0 and not allow_zero):
        fail(f"{name} must be {'non-negative' if allow_zero else 'positive'}")
    return v


def parse_shape(value: str) -> list[int]:
    parts = value.replace("X", "x").split("x")
    shape = []
    for p in parts:
        if not p or not p.lstrip("-").isdigit():
            fail(f"invalid --shape: {value}")
        n = int(p)
        if n <= 0:
            fail("shape dimensions must be positive")
        shape.append(n)
    if not 1 <= len(shape) <= 3:
        fail("--shape rank must be between 1 and 3")
    return shape


def parse_direction(value: str) -> str:
    if value in ("forward", "fwd", "fft"):
        return "forward"
    if value in ("inverse", "inv", "ifft"):
        return "inverse"
    fail(f"unknown direction: {value}")


def parse_placement(value: str) -> str:
    if value in ("in-place", "inplace", "in", "ip"):
        return "in-place"
    if value in ("out-of-place", "outofplace", "out", "oop"):
        return "out-of-place"
    fail(f"unknown placement: {value}")


def parse_args(argv: list[str]) -> dict:
    if not argv or argv[0] in ("-h", "--help"):
        print(USAGE, end="")
        raise SystemExit(EXIT_PASSED)
    args = {"command": argv[0], "rank": None, "shapes": [], "api": "c2c", "batch": 1, "direction": "forward",
            "placement": "out-of-place", "warmup": 10, "iters": 100, "json": False, "print_path": False,
            "ref": True, "check": False}
    if args["command"] == "tune":
        args["rest"] = argv[1:]
        return args
    if args["command"] != "bench":
        fail(f"unknown command: {args['command']}")
    i = 1

    def need(name):
        nonlocal i
        if i + 1 >= len(argv):
            fail(f"missing value for {name}")
        i += 1
        return argv[i]

    while i < len(argv):
        a = argv[i]
        if a in ("-h", "--help"):
            print(USAGE, end="")
            raise SystemExit(EXIT_PASSED)
        elif a == "--rank":
            v = need(a)
            if v not in ("1", "2", "3"):
                fail(f"invalid --rank: {v} (expected 1, 2, or 3)")
            args["rank"] = int(v)
        elif a == "--shape":
            args["shapes"] = [parse_shape(s) for s in need(a).split(",") if s]
        elif a == "--api":
            v = need(a).lower()
            if v not in APIS:
                fail(f"unknown FFT API: {v}")
            args["api"] = v
        elif a == "--batch":
            args["batch"] = parse_positive(a, need(a))
        elif a == "--direction":
            args["direction"] = parse_direction(need(a))
        elif a == "--placement":
            args["placement"] = parse_placement(need(a))
        elif a == "--warmup":
            args["warmup"] = parse_positive(a, need(a), allow_zero=True)
        elif a == "--iters":
            args["iters"] = parse_positive(a, need(a))
        elif a == "--json":
            args["json"] = True
        elif a == "--print-path":
            args["print_path"] = True
        elif a == "--no-ref":
            args["ref"] = False
        elif a == "--check":
            args["check"] = True
        else:
            fail(f"unknown argument: {a}")
        i += 1
    if args["rank"] is None:
        fail("--rank is required")
    if not args["shapes"]:
        fail("--shape is required")
    api = args["api"]
    for shape in args["shapes"]:
        if len(shape) != args["rank"]:
            fail("--shape dimension count does not match --rank")
        if args["rank"] == 2 and api not in COMPLEX and args["batch"] != 1:
            fail("--batch is only supported with --rank 1 or complex --rank 2")
    if api in REAL_FWD and args["direction"] != "forward":
        fail(f"{api} only supports forward direction")
    if api in REAL_INV and args["direction"] != "inverse":
        fail(f"{api} only supports inverse direction")
    return args


class Case:
    """One benchmark case: buffers, this library's plan and the comparator."""

    def __init__(self, spec: dict, torch, ff, dev):
        self.spec = spec
        self.torch, self.ff = torch, ff
        api, shape, batch = spec["api"], spec["shape"], spec["batch"]
        self.dbl = api in DOUBLE
        rdt = torch.float64 if self.dbl else torch.float32
        self.rdt, self.cdt = rdt, (torch.complex128 if self.dbl else torch.complex64)
        n = shape[-1]
        self.h = n // 2 + 1
        outer = batch * math.prod(shape[:-1])
        self.outer = outer
        padded = 2 * self.h
        in_sc = n if api in REAL_FWD else (padded if api in REAL_INV else 2 * n)
        out_sc = padded if api in REAL_FWD else (n if api in REAL_INV else 2 * n)
        inplace = spec["placement"] == "in-place"
        ip_sc = 2 * n if api in COMPLEX else padded
        count = outer * max(ip_sc if inplace else in_sc, ip_sc if inplace else out_sc)
        esz = 8 if self.dbl else 4
        # algorithmic bytes at the API boundary (compact rows)
        in_b = n * esz if api in REAL_FWD else (self.h * 2 * esz if api in REAL_INV else 2 * n * esz)
        out_b = self.h * 2 * esz if api in REAL_FWD else (n * esz if api in REAL_INV else 2 * n * esz)
        self.alg_bytes = outer * (in_b + out_b)
        T = ff.Type
        self.type = {"c2c": T.C2C, "z2z": T.Z2Z, "r2c": T.R2C, "d2z": T.D2Z, "c2r": T.C2R, "z2d": T.Z2D}[api]

        def seeded():
            buf = torch.sin(torch.arange(1, count + 1, dtype=rdt, device=dev) * 0.173)
            if api in REAL_INV:
                rows = buf[: outer * padded].view(outer, padded)
                rows[:, 1] = 0
                if n % 2 == 0:
                    rows[:, n + 1] = 0
            return buf

        self.ff_in = seeded()
        self.ff_out = self.ff_in if inplace else torch.zeros(count, dtype=rdt, device=dev)
        # plan: rank 1 compact rows out of place, padded real rows in place
        rank = spec["rank"]
        if rank == 1 and inplace and api not in COMPLEX:
            idist = padded if api in REAL_FWD else self.h
            odist = self.h if api in REAL_FWD else padded
            self.plan = ff.Plan.many(1, [n], None, 1, idist, None, 1, odist, self.type, batch)
            self.in_pitch, self.out_pitch = (padded, padded)
        elif rank == 1:
            self.plan = ff.Plan.one_d(n, self.type, batch)
            self.in_pitch = n if api in REAL_FWD else (padded if api in REAL_INV else 2 * n)
            self.out_pitch = padded if api in REAL_FWD else (n if api in REAL_INV else 2 * n)
        elif rank == 2 and api in COMPLEX:
            dist = shape[0] * shape[1]
            self.plan = ff.Plan.many(2, list(shape), None, 1, dist, None, 1, dist, self.type, batch)
            self.in_pitch = self.out_pitch = 2 * n
        elif rank == 2:
            self.plan = ff.Plan.two_d(shape[0], shape[1], self.type)
            self.in_pitch = n if api in REAL_FWD else padded
            self.out_pitch = padded if api in REAL_FWD else n
        else:  # rank 3 through the extension (flagfftPlan3d answers NOT_SUPPORTED, as in the reference)
            self.plan = ff.Plan.three_d(shape[0], shape[1], shape[2], self.type, batch)
            if api in COMPLEX:
                self.in_pitch = self.out_pitch = 2 * n
            else:
                self.in_pitch = n if api in REAL_FWD else padded
                self.out_pitch = padded if api in REAL_FWD else n
        fn = {"c2c": self.plan.c2c, "z2z": self.plan.z2z, "r2c": self.plan.r2c, "d2z": self.plan.d2z,
              "c2r": self.plan.c2r, "z2d": self.plan.z2d}[api]
        d = ff.INVERSE if spec["direction"] == "inverse" else ff.FORWARD
        i_ptr, o_ptr = self.ff_in.data_ptr(), self.ff_out.data_ptr()
        self.run_ff = (lambda: fn(i_ptr, o_ptr, d)) if api in COMPLEX else (lambda: fn(i_ptr, o_ptr))
        self.run_ref = self._ref(seeded(), dev) if spec["ref"] else None

    def _ref(self, buf, dev):
        """torch.fft on a second buffer of the same contents (unnormalised inverse, like the library)."""
        torch, spec = self.torch, self.spec
        api, shape = spec["api"], spec["shape"]
        n, h, outer = shape[-1], self.h, self.outer
        rank = spec["rank"]
        lead = [self.spec["batch"]] + list(shape[:-1])
        inv = spec["direction"] == "inverse"
        dims = tuple(range(-rank, 0))
        if api in COMPLEX:
            x = buf[: outer * 2 * n].view(self.cdt).view(*lead, n)
            out = torch.empty_like(x)
            f = torch.fft.ifftn if inv else torch.fft.fftn
            self.ref_out = out
            return lambda: f(x, dim=dims, norm="forward" if inv else "backward", out=out)
        if api in REAL_FWD:
            x = buf[: outer * self.in_pitch].view(outer, self.in_pitch)[:, :n].reshape(*lead, n)
            out = torch.empty(*lead, h, dtype=self.cdt, device=dev)
            self.ref_out = out
            return lambda: torch.fft.rfftn(x, dim=dims, out=out)
        X = buf[: outer * 2 * h].view(self.cdt).view(*lead, h)
        out = torch.empty(*lead, n, dtype=self.rdt, device=dev)
        s = list(shape)
        self.ref_out = out
        return lambda: torch.fft.irfftn(X, s=s, dim=dims, norm="forward", out=out)

    def ff_result(self):
        """This library's output as a tensor shaped like the comparator's."""
        torch, spec = self.torch, self.spec
        api, shape = spec["api"], spec["shape"]
        n, h, outer = shape[-1], self.h, self.outer
        lead = [spec["batch"]] + list(shape[:-1])
        o = self.ff_out
        if api in COMPLEX:
            return o[: outer * 2 * n].view(self.cdt).view(*lead, n)
        if api in REAL_FWD:
            return o[: outer * self.out_pitch].view(outer, self.out_pitch)[:, : 2 * h].contiguous().view(
                self.cdt).view(*lead, h)
        return o[: outer * self.out_pitch].view(outer, self.out_pitch)[:, :n].reshape(*lead, n)


def percentile(times: list[float], fraction: float) -> float:
    s = sorted(times)
    return s[min(len(s) - 1, int(fraction * len(s)))]


def run_case(spec: dict, warmup: int, iters: int, torch, ff, dev) -> dict:
    c = Case(spec, torch, ff, dev)
    stream = torch.cuda.Stream(device=dev)
    c.plan.set_stream(stream.cuda_stream)
    ev = [torch.cuda.Event(enable_timing=True) for _ in range(2)]

    def timed(fn):
        ev[0].record(stream)
        fn()
        ev[1].record(stream)
        ev[1].synchronize()
        return ev[0].elapsed_time(ev[1])

    ff_t, ref_t = [], []
    with torch.cuda.stream(stream):
        for _ in range(warmup):
            c.run_ff()
            if c.run_ref:
                c.run_ref()
        torch.cuda.synchronize()
        for i in range(iters):
            order = [(c.run_ff, ff_t), (c.run_ref, ref_t)] if i % 2 else [(c.run_ref, ref_t), (c.run_ff, ff_t)]
            for fn, acc in order:
                if fn is not None:
                    acc.append(timed(fn))
        torch.cuda.synchronize()
    res = {"ff": (percentile(ff_t, 0.5), percentile(ff_t, 0.9))}
    res["ref"] = (percentile(ref_t, 0.5), percentile(ref_t, 0.9)) if ref_t else (0.0, 0.0)
    res["speedup"] = res["ref"][0] / res["ff"][0] if ref_t and res["ff"][0] > 0 else 0.0
    d = ff.INVERSE if spec["direction"] == "inverse" else ff.FORWARD
    r, launches, pass_bytes = ff.plan_stats(c.plan.handle, d)
    ff.check("flagfftxPlanStats", r)
    res["bytes"] = {"algorithmic": c.alg_bytes, "launches": launches, "plan_pass_bytes": pass_bytes,
                    "flagfft_gbps": c.alg_bytes / (res["ff"][0] * 1e6),
                    "ref_gbps": c.alg_bytes / (res["ref"][0] * 1e6) if ref_t else None}
    if spec["check"] and c.run_ref is not None:
        with torch.cuda.stream(stream):
            if spec["placement"] == "in-place":
                # one more exec on the freshly seeded input
                fresh = Case(spec, torch, ff, dev)
                fresh.plan.set_stream(stream.cuda_stream)
                fresh.run_ff()
                c = fresh
                c.run_ref()
            else:
                c.run_ff()
                c.run_ref()
            torch.cuda.synchronize()
        a = c.ff_result().reshape(-1).to(torch.complex128)
        b = c.ref_out.reshape(-1).to(torch.complex128)
        num = torch.linalg.vector_norm(a - b).item()
        den = torch.linalg.vector_norm(b).item()
        res["rel_l2_vs_ref"] = num / den if den > 0 else num
    if spec["print_path"]:
        res["plan_description"] = c.plan.description()
    return res


def case_json(spec: dict) -> dict:
    return {"api": spec["api"], "shape": spec["shape"], "rank": spec["rank"], "batch": spec["batch"],
            "direction": spec["direction"], "placement": spec["placement"]}


def format_json(specs, results, warmup, iters) -> dict:
    cases = []
    for spec, r in zip(specs, results):
        e = case_json(spec)
        e["timing"] = {"flagfft_median_ms": r["ff"][0], "flagfft_p90_ms": r["ff"][1], "ref_median_ms": r["ref"][0],
                       "ref_p90_ms": r["ref"][1], "speedup": r["speedup"], "warmup": warmup, "iters": iters}
        e["ref"] = "torch.fft (hipFFT / rocFFT)" if spec["ref"] else None
        e["bytes"] = r["bytes"]
        if "rel_l2_vs_ref" in r:
            e["rel_l2_vs_ref"] = r["rel_l2_vs_ref"]
        if r.get("plan_description"):
            e["plan_description"] = r["plan_description"]
        cases.append(e)
    return {"status": "passed", "command": "bench", "cases": cases}


def format_table(specs, results, warmup, iters) -> str:
    head = (f"{'shape':<12}{'api':<8}{'batch':<7}{'direction':<11}{'placement':<14}{'warmup':<8}{'iters':<7}"
            f"{'flagfft_median_ms':<19}{'ref_median_ms':<15}{'speedup':<9}flagfft_GBps\n")
    out = [head]
    for spec, r in zip(specs, results):
        shape = "x".join(str(s) for s in spec["shape"])
        out.append(f"{shape:<12}{spec['api']:<8}{spec['batch']:<7}{spec['direction']:<11}{spec['placement']:<14}"
                   f"{warmup:<8}{iters:<7}{r['ff'][0]:<19.4f}{r['ref'][0]:<15.4f}{r['speedup']:<9.2f}"
                   f"{r['bytes']['flagfft_gbps']:.0f}\n")
        if r.get("plan_description"):
            out.append(r["plan_description"] + "\n")
    return "".join(out)


def main(argv: list[str] | None = None) -> int:
    argv = sys.argv[1:] if argv is None else argv
    try:
        args = parse_args(argv)
        if args["command"] == "tune":
            from . import autotune

            sys.argv = ["flagfft_amd.autotune"] + args["rest"]
            return autotune.main() or 0
        import torch

        if not torch.cuda.is_available():
            print("error: no HIP device available", file=sys.stderr)
            return EXIT_SKIPPED
        from . import flagfft as ff

        dev = torch.device("cuda:0")
        specs = []
        for shape in args["shapes"]:
            s = {k: args[k] for k in ("rank", "api", "batch", "direction", "placement", "ref", "check",
                                      "print_path")}
            s["shape"] = shape
            specs.append(s)
        results = [run_case(s, args["warmup"], args["iters"], torch, ff, dev) for s in specs]
        if args["json"]:
            print(json.dumps(format_json(specs, results, args["warmup"], args["iters"]), indent=2))
        else:
            print(format_table(specs, results, args["warmup"], args["iters"]), end="")
        return EXIT_PASSED
    except CliError as e:
        print(f"error: {e}", file=sys.stderr)
        return e.code
    except SystemExit as e:
        return int(e.code or 0)
    except Exception as e:  # library errors (FlagFFTError) and device failures
        print(f"error: {e}", file=sys.stderr)
        return EXIT_RUNTIME_ERROR


if __name__ == "__main__":
    sys.exit(main())
