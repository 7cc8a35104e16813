"""ctypes mirror of the C API in include/flagfft.h (and include/flagfft_ext.h).

Same names, argument meaning and result codes as the reference's C API
(include/flagfft.h:1-86 of the reference); device buffers are passed as raw
device pointers (ints), e.g. ``tensor.data_ptr()`` of a torch tensor on the GPU.
The library is the in-tree libflagfft.so; importing this module raises if it
has not been built.
"""
from __future__ import annotations

import ctypes
import enum
import os

_PKG = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_PKG, "libflagfft.so")


class Result(enum.IntEnum):
    SUCCESS = 0
    INVALID_PLAN = 1
    ALLOC_FAILED = 2
    INVALID_TYPE = 3
    INVALID_VALUE = 4
    INTERNAL_ERROR = 5
    EXEC_FAILED = 6
    SETUP_FAILED = 7
    INVALID_SIZE = 8
    UNALIGNED_DATA = 9
    INCOMPLETE_PARAMETER_LIST = 10
    INVALID_DEVICE = 11
    PARSE_ERROR = 12
    NO_WORKSPACE = 13
    NOT_SUPPORTED = 14


class Type(enum.IntEnum):
    R2C = 0x2A
    C2R = 0x2C
    C2C = 0x29
    D2Z = 0x6A
    Z2D = 0x6C
    Z2Z = 0x69


FORWARD = -1
INVERSE = 1

SYMBOLS = [
    "flagfftPlan1d", "flagfftPlan2d", "flagfftPlan3d", "flagfftPlanMany",
    "flagfftExecC2C", "flagfftExecZ2Z", "flagfftExecR2C", "flagfftExecD2Z", "flagfftExecC2R", "flagfftExecZ2D",
    "flagfftSetStream", "flagfftDestroy", "flagfftGetPlanDescription",
]
EXT_SYMBOLS = ["flagfftxVersion", "flagfftxPlanDescribe", "flagfftxPlanStats", "flagfftxSetTuning",
               "flagfftxSetProfiling", "flagfftxGetProfile", "flagfftxStepLabel", "flagfftxPlan1dCompute", "flagfftxPlan3d", "flagfftxPlanMany",
               "flagfftxSetGraphMode", "flagfftxServerCreate", "flagfftxServerPrepare", "flagfftxServerSubmit",
               "flagfftxServerPoll", "flagfftxServerWait", "flagfftxServerRelease", "flagfftxServerGetStats",
               "flagfftxServerResetStats", "flagfftxServerDestroy", "flagfftxTunedStore", "flagfftxTunedDbPath",
               "flagfftxExecC2CNoAlloc"]

SERVE_BATCHED, SERVE_PLAIN = 0, 1

# flagfftxCompute (include/flagfft_ext.h)
COMPUTE_VECTOR, COMPUTE_BF16X3, COMPUTE_FP16X3 = 0, 1, 2


class FlagFFTError(RuntimeError):
    def __init__(self, fn: str, code: int):
        try:
            name = Result(code).name
        except ValueError:
            name = str(code)
        super().__init__(f"{fn} failed with FLAGFFT_{name}")
        self.code = code


class Tuning(ctypes.Structure):
    _fields_ = [
        ("leaf_max", ctypes.c_int), ("pass_max", ctypes.c_int), ("units", ctypes.c_int), ("ept", ctypes.c_int),
        ("chunk_bytes", ctypes.c_int), ("npasses", ctypes.c_int), ("passes", ctypes.c_int * 4),
        ("nradices", ctypes.c_int), ("radices", ctypes.c_int * 16), ("stage_form", ctypes.c_int),
        ("variant", ctypes.c_int),
    ]


class ServerStats(ctypes.Structure):
    _fields_ = [
        ("completed", ctypes.c_longlong), ("transforms", ctypes.c_longlong), ("bytes", ctypes.c_double),
        ("span_s", ctypes.c_double), ("p50_ms", ctypes.c_double), ("p99_ms", ctypes.c_double),
        ("mean_ms", ctypes.c_double), ("max_ms", ctypes.c_double), ("batches", ctypes.c_longlong),
        ("queue_launches", ctypes.c_longlong), ("queued_requests", ctypes.c_longlong),
        ("arena_peak_bytes", ctypes.c_double), ("table_bytes", ctypes.c_double),
    ]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


_lib = None
_H = ctypes.c_void_p
_HP = ctypes.POINTER(ctypes.c_void_p)
_IP = ctypes.POINTER(ctypes.c_int)
_int = ctypes.c_int
_vp = ctypes.c_void_p


def lib():
    """Load libflagfft.so (raises FileNotFoundError when it is not built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise FileNotFoundError(f"{LIB_PATH} is not built; run __graft_entry__.build()")
        L = ctypes.CDLL(LIB_PATH)
        sigs = {
            "flagfftPlan1d": [_HP, _int, _int, _int],
            "flagfftPlan2d": [_HP, _int, _int, _int],
            "flagfftPlan3d": [_HP, _int, _int, _int, _int],
            "flagfftPlanMany": [_HP, _int, _IP, _IP, _int, _int, _IP, _int, _int, _int, _int],
            "flagfftExecC2C": [_H, _vp, _vp, _int],
            "flagfftxExecC2CNoAlloc": [_H, _vp, _vp, _int],
            "flagfftExecZ2Z": [_H, _vp, _vp, _int],
            "flagfftExecR2C": [_H, _vp, _vp],
            "flagfftExecD2Z": [_H, _vp, _vp],
            "flagfftExecC2R": [_H, _vp, _vp],
            "flagfftExecZ2D": [_H, _vp, _vp],
            "flagfftSetStream": [_H, _vp],
            "flagfftDestroy": [_H],
            "flagfftxPlanDescribe": [_int, _IP, _int, _int, ctypes.c_char_p, ctypes.c_size_t],
            "flagfftxPlanStats": [_H, _int, _IP, ctypes.POINTER(ctypes.c_longlong)],
            "flagfftxPlan1dCompute": [_HP, _int, _int, _int, _int],
            "flagfftxPlan3d": [_HP, _int, _int, _int, _int, _int],
            "flagfftxPlanMany": [_HP, _int, _IP, _IP, _int, _int, _IP, _int, _int, _int, _int],
            "flagfftxSetGraphMode": [_H, _int],
            "flagfftxTunedStore": [_int, _int, _int, ctypes.POINTER(Tuning), ctypes.c_double, ctypes.c_double,
                                   ctypes.c_double],
            "flagfftxTunedDbPath": [ctypes.c_char_p, ctypes.c_size_t],
            "flagfftxServerCreate": [_HP, ctypes.c_size_t, _int, _int],
            "flagfftxServerPrepare": [_H, _int, _int, _int],
            "flagfftxServerSubmit": [_H, _int, _int, _int, _int, _vp, ctypes.POINTER(ctypes.c_longlong)],
            "flagfftxServerPoll": [_H, ctypes.c_longlong, _IP, ctypes.POINTER(ctypes.c_void_p),
                                   ctypes.POINTER(ctypes.c_double)],
            "flagfftxServerWait": [_H, ctypes.c_longlong, ctypes.POINTER(ctypes.c_void_p),
                                   ctypes.POINTER(ctypes.c_double)],
            "flagfftxServerRelease": [_H, ctypes.c_longlong],
            "flagfftxServerGetStats": [_H, ctypes.POINTER(ServerStats)],
            "flagfftxServerResetStats": [_H],
            "flagfftxServerDestroy": [_H],
            "flagfftxSetTuning": [ctypes.POINTER(Tuning)],
            "flagfftxSetProfiling": [_H, _int],
            "flagfftxGetProfile": [_H, _int, _int, ctypes.POINTER(ctypes.c_double),
                                   ctypes.POINTER(ctypes.c_longlong), ctypes.POINTER(ctypes.c_longlong), _IP],
        }
        for name, argtypes in sigs.items():
            fn = getattr(L, name)
            fn.argtypes = argtypes
            fn.restype = _int
        L.flagfftGetPlanDescription.argtypes = [_H]
        L.flagfftGetPlanDescription.restype = ctypes.c_char_p
        L.flagfftxStepLabel.argtypes = [_H, _int, _int]
        L.flagfftxStepLabel.restype = ctypes.c_char_p
        L.flagfftxVersion.argtypes = []
        L.flagfftxVersion.restype = ctypes.c_char_p
        _lib = L
    return _lib


def _arr(values):
    if values is None:
        return None
    return (ctypes.c_int * len(values))(*values)


# ---- raw calls returning result codes (mirror the C signatures) ------------
def plan1d(nx: int, type_: int, batch: int):
    h = ctypes.c_void_p()
    r = lib().flagfftPlan1d(ctypes.byref(h), nx, int(type_), batch)
    return r, h


def plan1d_compute(nx: int, type_: int, batch: int, compute: int):
    h = ctypes.c_void_p()
    r = lib().flagfftxPlan1dCompute(ctypes.byref(h), nx, int(type_), batch, int(compute))
    return r, h


def plan3d_ext(nx: int, ny: int, nz: int, type_: int, batch: int = 1):
    h = ctypes.c_void_p()
    r = lib().flagfftxPlan3d(ctypes.byref(h), nx, ny, nz, int(type_), batch)
    return r, h


def plan2d(nx: int, ny: int, type_: int):
    h = ctypes.c_void_p()
    r = lib().flagfftPlan2d(ctypes.byref(h), nx, ny, int(type_))
    return r, h


def plan3d(nx: int, ny: int, nz: int, type_: int):
    h = ctypes.c_void_p()
    r = lib().flagfftPlan3d(ctypes.byref(h), nx, ny, nz, int(type_))
    return r, h


def plan_many(rank, n, inembed, istride, idist, onembed, ostride, odist, type_, batch):
    h = ctypes.c_void_p()
    r = lib().flagfftPlanMany(ctypes.byref(h), rank, _arr(n), _arr(inembed), istride, idist, _arr(onembed), ostride,
                              odist, int(type_), batch)
    return r, h


def plan_many_ext(rank, n, inembed, istride, idist, onembed, ostride, odist, type_, batch):
    h = ctypes.c_void_p()
    r = lib().flagfftxPlanMany(ctypes.byref(h), rank, _arr(n), _arr(inembed), istride, idist, _arr(onembed), ostride,
                               odist, int(type_), batch)
    return r, h


def exec_c2c(h, idata: int, odata: int, direction: int) -> int:
    return lib().flagfftExecC2C(h, idata, odata, direction)


def exec_c2c_noalloc(h, idata: int, odata: int, direction: int) -> int:
    """flagfftxExecC2CNoAlloc: flagfftExecC2C with the workspaces in the library's fixed per-device array."""
    return lib().flagfftxExecC2CNoAlloc(h, idata, odata, direction)


def exec_z2z(h, idata: int, odata: int, direction: int) -> int:
    return lib().flagfftExecZ2Z(h, idata, odata, direction)


def exec_r2c(h, idata: int, odata: int) -> int:
    return lib().flagfftExecR2C(h, idata, odata)


def exec_d2z(h, idata: int, odata: int) -> int:
    return lib().flagfftExecD2Z(h, idata, odata)


def exec_c2r(h, idata: int, odata: int) -> int:
    return lib().flagfftExecC2R(h, idata, odata)


def exec_z2d(h, idata: int, odata: int) -> int:
    return lib().flagfftExecZ2D(h, idata, odata)


def set_stream(h, stream: int) -> int:
    return lib().flagfftSetStream(h, stream)


def destroy(h) -> int:
    return lib().flagfftDestroy(h)


def description(h) -> str | None:
    s = lib().flagfftGetPlanDescription(h)
    return s.decode() if s is not None else None


def describe_host(rank: int, n, type_: int, batch: int = 1):
    """Planner-only description (no device is touched): (code, text)."""
    buf = ctypes.create_string_buffer(1 << 16)
    r = lib().flagfftxPlanDescribe(rank, _arr(n), int(type_), batch, buf, len(buf))
    return r, buf.value.decode()


def plan_stats(h, direction: int = FORWARD):
    launches = ctypes.c_int(0)
    nbytes = ctypes.c_longlong(0)
    r = lib().flagfftxPlanStats(h, direction, ctypes.byref(launches), ctypes.byref(nbytes))
    return r, launches.value, nbytes.value


def make_tuning(**kw) -> Tuning:
    """flagfftxTuning from keyword overrides (leaf_max, pass_max, units, ept, chunk_bytes, passes, radices,
    stage_form, variant)."""
    t = Tuning()
    for k, v in kw.items():
        if k == "passes":
            t.npasses = len(v)
            for i, x in enumerate(v):
                t.passes[i] = x
        elif k == "radices":
            t.nradices = len(v)
            for i, x in enumerate(v):
                t.radices[i] = x
        else:
            setattr(t, k, int(v))
    return t


def set_tuning(**kw) -> int:
    """Plan-time tunables for plans created afterwards on this thread; no args restores defaults."""
    if not kw:
        return lib().flagfftxSetTuning(None)
    t = make_tuning(**kw)
    return lib().flagfftxSetTuning(ctypes.byref(t))


def tuned_store(n: int, type_: int, batch: int, median_us: float, gbps: float, default_us: float, **kw) -> int:
    """Append a tuned-plan row (the overrides in kw) to the tuned-plan database."""
    t = make_tuning(**kw)
    return lib().flagfftxTunedStore(n, int(type_), batch, ctypes.byref(t), median_us, gbps, default_us)


def tuned_db_path() -> str:
    buf = ctypes.create_string_buffer(4096)
    check("flagfftxTunedDbPath", lib().flagfftxTunedDbPath(buf, len(buf)))
    return buf.value.decode()

def set_profiling(h, enable: bool) -> int:
    return lib().flagfftxSetProfiling(h, 1 if enable else 0)


def get_profile(h, direction: int = FORWARD):
    """Per-launch [(label, total_ms, calls, bytes)] of one Exec direction (waits for the events)."""
    n = ctypes.c_int(0)
    r = lib().flagfftxGetProfile(h, direction, 0, None, None, None, ctypes.byref(n))
    check("flagfftxGetProfile", r)
    cap = n.value
    ms = (ctypes.c_double * max(cap, 1))()
    calls = (ctypes.c_longlong * max(cap, 1))()
    nbytes = (ctypes.c_longlong * max(cap, 1))()
    r = lib().flagfftxGetProfile(h, direction, cap, ms, calls, nbytes, ctypes.byref(n))
    check("flagfftxGetProfile", r)
    out = []
    for i in range(cap):
        label = lib().flagfftxStepLabel(h, direction, i)
        out.append((label.decode() if label else "", ms[i], calls[i], nbytes[i]))
    return out


def version() -> str:
    return lib().flagfftxVersion().decode()


def check(fn: str, code: int):
    if code != Result.SUCCESS:
        raise FlagFFTError(fn, code)


# ---- convenience object --------------------------------------------------
class Plan:
    """Owning wrapper of a flagfftHandle (raises FlagFFTError on failure)."""

    def __init__(self, handle):
        self.handle = handle

    @classmethod
    def one_d(cls, n: int, type_: int, batch: int = 1):
        r, h = plan1d(n, type_, batch)
        check("flagfftPlan1d", r)
        return cls(h)

    @classmethod
    def one_d_compute(cls, n: int, type_: int, batch: int, compute: int):
        r, h = plan1d_compute(n, type_, batch, compute)
        check("flagfftxPlan1dCompute", r)
        return cls(h)

    @classmethod
    def two_d(cls, nx: int, ny: int, type_: int):
        r, h = plan2d(nx, ny, type_)
        check("flagfftPlan2d", r)
        return cls(h)

    @classmethod
    def three_d(cls, nx: int, ny: int, nz: int, type_: int, batch: int = 1):
        r, h = plan3d_ext(nx, ny, nz, type_, batch)
        check("flagfftxPlan3d", r)
        return cls(h)

    @classmethod
    def many_ext(cls, rank, n, inembed, istride, idist, onembed, ostride, odist, type_, batch):
        r, h = plan_many_ext(rank, n, inembed, istride, idist, onembed, ostride, odist, type_, batch)
        check("flagfftxPlanMany", r)
        return cls(h)

    @classmethod
    def many(cls, rank, n, inembed, istride, idist, onembed, ostride, odist, type_, batch):
        r, h = plan_many(rank, n, inembed, istride, idist, onembed, ostride, odist, type_, batch)
        check("flagfftPlanMany", r)
        return cls(h)

    def set_graph_mode(self, enable: bool = True):
        check("flagfftxSetGraphMode", lib().flagfftxSetGraphMode(self.handle, int(bool(enable))))

    def set_stream(self, stream: int):
        check("flagfftSetStream", set_stream(self.handle, stream))

    def c2c(self, i: int, o: int, direction: int = FORWARD):
        check("flagfftExecC2C", exec_c2c(self.handle, i, o, direction))

    def z2z(self, i: int, o: int, direction: int = FORWARD):
        check("flagfftExecZ2Z", exec_z2z(self.handle, i, o, direction))

    def r2c(self, i: int, o: int):
        check("flagfftExecR2C", exec_r2c(self.handle, i, o))

    def d2z(self, i: int, o: int):
        check("flagfftExecD2Z", exec_d2z(self.handle, i, o))

    def c2r(self, i: int, o: int):
        check("flagfftExecC2R", exec_c2r(self.handle, i, o))

    def z2d(self, i: int, o: int):
        check("flagfftExecZ2D", exec_z2d(self.handle, i, o))

    def description(self) -> str:
        return description(self.handle)

    def stats(self, direction: int = FORWARD):
        r, launches, nbytes = plan_stats(self.handle, direction)
        check("flagfftxPlanStats", r)
        return launches, nbytes

    def destroy(self):
        if self.handle is not None and self.handle.value:
            check("flagfftDestroy", destroy(self.handle))
        self.handle = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.destroy()

    def __del__(self):
        try:
            if self.handle is not None and self.handle.value:
                destroy(self.handle)
        except Exception:
            pass


class Server:
    """Owning wrapper of a flagfftxServer (the serving loop of include/flagfft_ext.h)."""

    def __init__(self, budget_bytes: int, max_inflight: int = 3, mode: int = SERVE_BATCHED):
        h = ctypes.c_void_p()
        check("flagfftxServerCreate", lib().flagfftxServerCreate(ctypes.byref(h), budget_bytes, max_inflight, mode))
        self.handle = h

    def prepare(self, type_: int, n: int, direction: int = FORWARD):
        check("flagfftxServerPrepare", lib().flagfftxServerPrepare(self.handle, int(type_), n, direction))

    def submit(self, type_: int, n: int, batch: int, idata: int, direction: int = FORWARD) -> int:
        t = ctypes.c_longlong()
        check("flagfftxServerSubmit",
              lib().flagfftxServerSubmit(self.handle, int(type_), n, batch, direction, idata, ctypes.byref(t)))
        return t.value

    def poll(self, ticket: int):
        """(done, output device pointer or None, latency ms)."""
        d = ctypes.c_int()
        o = ctypes.c_void_p()
        l = ctypes.c_double()
        check("flagfftxServerPoll", lib().flagfftxServerPoll(self.handle, ticket, ctypes.byref(d), ctypes.byref(o),
                                                            ctypes.byref(l)))
        return bool(d.value), o.value, l.value

    def wait(self, ticket: int):
        """(output device pointer, latency ms), blocking."""
        o = ctypes.c_void_p()
        l = ctypes.c_double()
        check("flagfftxServerWait", lib().flagfftxServerWait(self.handle, ticket, ctypes.byref(o), ctypes.byref(l)))
        return o.value, l.value

    def release(self, ticket: int):
        check("flagfftxServerRelease", lib().flagfftxServerRelease(self.handle, ticket))

    def stats(self) -> dict:
        s = ServerStats()
        check("flagfftxServerGetStats", lib().flagfftxServerGetStats(self.handle, ctypes.byref(s)))
        return s.as_dict()

    def reset_stats(self):
        check("flagfftxServerResetStats", lib().flagfftxServerResetStats(self.handle))

    def close(self):
        if self.handle:
            r = lib().flagfftxServerDestroy(self.handle)
            self.handle = None
            check("flagfftxServerDestroy", r)

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()
        return False
