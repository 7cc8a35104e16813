"""C-ABI boundary and host logic, no GPU needed.

* libflagfft.so loads and exports every function declared in include/*.h;
* descriptor validation returns the reference's codes before touching a device
  (ctest/test_plan.cpp; src/exec/c_api_plan.cpp:274-336; c_api_internal.cpp:60-120);
* the planner (flagfftxPlanDescribe, host only) produces decompositions whose
  integer invariants hold (products, radix set, LDS budget, Rader/Bluestein choice).
"""
import math
import os
import re

import pytest

from flagfft_amd import flagfft as ff

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
R = ff.Result
T = ff.Type
ALL_TYPES = [T.C2C, T.Z2Z, T.R2C, T.D2Z, T.C2R, T.Z2D]


def _declared_functions():
    names = []
    for h in ("flagfft.h", "flagfft_ext.h"):
        text = open(os.path.join(ROOT, "include", h)).read()
        text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
        names += re.findall(r"^\s*(?:const\s+)?[A-Za-z_][\w\s\*]*?\b(flagfft\w+)\s*\(", text, flags=re.M)
    return sorted(set(names))


def test_library_exports_every_declared_symbol():
    L = ff.lib()
    declared = _declared_functions()
    assert set(ff.SYMBOLS) <= set(declared) and set(ff.EXT_SYMBOLS) <= set(declared)
    for name in declared:
        assert hasattr(L, name), name
    assert ff.version().startswith("flagfft-gfx950")


def test_enum_values_match_cufft_layout():
    """include/flagfft.h:28-35 of the reference uses cuFFT's values."""
    assert (T.R2C, T.C2R, T.C2C, T.D2Z, T.Z2D, T.Z2Z) == (0x2A, 0x2C, 0x29, 0x6A, 0x6C, 0x69)
    assert ff.FORWARD == -1 and ff.INVERSE == 1
    assert [r.value for r in R] == list(range(15))


def test_plan_invalid_parameters():
    """ctest/test_plan.cpp:26-36, 126-131, 155-169."""
    r, h = ff.plan1d(0, T.C2C, 1)
    assert r == R.INVALID_SIZE and not h.value
    assert ff.lib().flagfftPlan1d(None, 256, int(T.C2C), 1) == R.INVALID_VALUE
    r, h = ff.plan1d(256, T.C2C, 0)
    assert r == R.INVALID_VALUE and not h.value
    r, h = ff.plan2d(0, 32, T.C2C)
    assert r == R.INVALID_SIZE and not h.value
    assert ff.lib().flagfftPlan2d(None, 64, 32, int(T.C2C)) == R.INVALID_VALUE
    for t in ALL_TYPES:
        r, h = ff.plan3d(32, 16, 8, t)
        assert r == R.NOT_SUPPORTED and not h.value
    r, h = ff.plan3d(0, 16, 8, T.C2C)
    assert r == R.INVALID_SIZE
    assert ff.lib().flagfftPlan3d(None, 32, 16, 8, int(T.C2C)) == R.INVALID_VALUE
    r, h = ff.plan1d(256, 0x33, 1)
    assert r == R.INVALID_TYPE and not h.value
    r, h = ff.plan_many(0, [16], None, 1, 16, None, 1, 16, T.C2C, 1)
    assert r == R.INVALID_VALUE


def test_layouts_outside_the_contract_are_not_supported():
    """2D istride != 1 and padded 2D dist (ctest/test_plan.cpp:105-124); 1D strides."""
    n = [64, 32]
    logical = 64 * 32
    assert ff.plan_many(2, n, None, 2, logical, None, 1, logical, T.C2C, 1)[0] == R.NOT_SUPPORTED
    assert ff.plan_many(2, n, None, 1, logical + 16, None, 1, logical, T.C2C, 1)[0] == R.NOT_SUPPORTED
    assert ff.plan_many(2, n, None, 1, logical, None, 1, logical + 16, T.C2C, 1)[0] == R.NOT_SUPPORTED
    assert ff.plan_many(1, [256], None, 2, 256, None, 1, 256, T.C2C, 1)[0] == R.NOT_SUPPORTED
    assert ff.plan_many(1, [256], [256], 1, 300, [256], 1, 256, T.C2C, 2)[0] == R.NOT_SUPPORTED


def test_valid_descriptors_reach_the_device_check():
    """Without a GPU a valid descriptor passes validation and stops at INVALID_DEVICE."""
    import torch

    if torch.cuda.is_available():
        pytest.skip("a GPU is visible")
    for t in ALL_TYPES:
        assert ff.plan1d(256, t, 1)[0] == R.INVALID_DEVICE
    # padded in-place real layouts (src/exec/c_api_internal.cpp:74-89)
    assert ff.plan_many(1, [256], None, 1, 258, None, 1, 129, T.R2C, 4)[0] == R.INVALID_DEVICE
    assert ff.plan_many(1, [256], None, 1, 129, None, 1, 258, T.C2R, 4)[0] == R.INVALID_DEVICE
    assert ff.plan2d(64, 32, T.R2C)[0] == R.INVALID_DEVICE


def test_rank3_extension_validation():
    """flagfftxPlan3d argument checks; flagfftPlan3d itself keeps FLAGFFT_NOT_SUPPORTED."""
    import torch

    assert ff.lib().flagfftxPlan3d(None, 8, 8, 8, int(T.C2C), 1) == R.INVALID_VALUE
    assert ff.plan3d_ext(8, 8, 8, T.C2C, 0)[0] == R.INVALID_VALUE
    assert ff.plan3d_ext(0, 8, 8, T.C2C, 1)[0] == R.INVALID_SIZE
    assert ff.plan3d_ext(8, 8, -1, T.R2C, 1)[0] == R.INVALID_SIZE
    assert ff.plan3d_ext(8, 8, 8, 0x77, 1)[0] == R.INVALID_TYPE
    r, h = ff.plan3d(32, 16, 8, T.C2C)
    assert r == R.NOT_SUPPORTED and not h.value
    if not torch.cuda.is_available():
        for t in ALL_TYPES:
            r, h = ff.plan3d_ext(16, 12, 10, t, 2)
            assert r == R.INVALID_DEVICE and not h.value


def test_advanced_layout_extension_validation():
    """flagfftxPlanMany: general C2C / Z2Z layouts of rank 1-2 pass validation; flagfftPlanMany keeps
    NOT_SUPPORTED for them (ctest/test_plan.cpp:105-124); real layouts and rank 3 stay unsupported."""
    import torch

    strided = (2, [64, 32], [64, 35], 2, 64 * 35 * 2, [64, 40], 1, 64 * 40, T.C2C, 2)
    assert ff.plan_many(*strided)[0] == R.NOT_SUPPORTED
    assert ff.plan_many_ext(1, [256], None, 1, 256, None, 1, 256, 0x77, 1)[0] == R.INVALID_TYPE
    assert ff.plan_many_ext(1, [0], None, 1, 256, None, 1, 256, T.C2C, 1)[0] == R.INVALID_SIZE
    assert ff.plan_many_ext(1, [256], None, 0, 256, None, 1, 256, T.C2C, 1)[0] == R.INVALID_VALUE
    assert ff.plan_many_ext(1, [256], None, 2, 0, None, 1, 256, T.C2C, 1)[0] == R.INVALID_VALUE
    assert ff.plan_many_ext(2, [64, 32], [64, 31], 1, 4096, None, 1, 2048, T.C2C, 1)[0] == R.INVALID_VALUE
    assert ff.plan_many_ext(1, [256], None, 2, 600, None, 1, 256, T.R2C, 1)[0] == R.NOT_SUPPORTED
    # real 2D: padded rows with unit strides only, batches n0 rows apart, even real pitch
    assert ff.plan_many_ext(2, [16, 32], [16, 34], 2, 16 * 34, [16, 17], 1, 16 * 17, T.R2C, 1)[0] == R.NOT_SUPPORTED
    assert ff.plan_many_ext(2, [16, 32], [16, 34], 1, 16 * 34 + 1, [16, 17], 1, 16 * 17, T.R2C, 1)[0] == \
        R.NOT_SUPPORTED
    assert ff.plan_many_ext(2, [16, 32], [16, 35], 1, 16 * 35, [16, 17], 1, 16 * 17, T.R2C, 1)[0] == R.NOT_SUPPORTED
    assert ff.plan_many(2, [16, 32], [16, 34], 1, 16 * 34, [16, 17], 1, 16 * 17, T.R2C, 1)[0] == R.NOT_SUPPORTED
    assert ff.plan_many_ext(3, [8, 8, 8], None, 2, 2048, None, 1, 512, T.C2C, 1)[0] == R.NOT_SUPPORTED
    assert ff.lib().flagfftxPlanMany(None, 1, None, None, 1, 1, None, 1, 1, int(T.C2C), 1) == R.INVALID_VALUE
    if not torch.cuda.is_available():
        assert ff.plan_many_ext(*strided)[0] == R.INVALID_DEVICE
        assert ff.plan_many_ext(1, [256], None, 3, 800, None, 2, 600, T.Z2Z, 4)[0] == R.INVALID_DEVICE
        assert ff.plan_many_ext(2, [16, 32], [16, 34], 1, 16 * 34, [16, 17], 1, 16 * 17, T.R2C, 2)[0] == \
            R.INVALID_DEVICE
        assert ff.plan_many_ext(2, [16, 32], [16, 17], 1, 16 * 17, [16, 34], 1, 16 * 34, T.C2R, 2)[0] == \
            R.INVALID_DEVICE


def test_compute_mode_validation():
    """flagfftxPlan1dCompute: unknown mode, unsupported type / length, device check."""
    import torch

    assert ff.lib().flagfftxPlan1dCompute(None, 1024, int(T.C2C), 1, 1) == R.INVALID_VALUE
    assert ff.plan1d_compute(1024, T.C2C, 1, 3)[0] == R.INVALID_VALUE
    assert ff.plan1d_compute(1024, T.C2C, 1, -1)[0] == R.INVALID_VALUE
    for compute in (ff.COMPUTE_BF16X3, ff.COMPUTE_FP16X3):
        assert ff.plan1d_compute(2048, T.C2C, 1, compute)[0] == R.NOT_SUPPORTED
        assert ff.plan1d_compute(1024, T.Z2Z, 1, compute)[0] == R.NOT_SUPPORTED
        assert ff.plan1d_compute(1024, T.R2C, 1, compute)[0] == R.NOT_SUPPORTED
        assert ff.plan1d_compute(0, T.C2C, 1, compute)[0] == R.INVALID_SIZE
        assert ff.plan1d_compute(1024, T.C2C, 0, compute)[0] == R.INVALID_VALUE
        if not torch.cuda.is_available():
            r, h = ff.plan1d_compute(1024, T.C2C, 8, compute)
            assert r == R.INVALID_DEVICE and not h.value


def test_server_validation():
    """flagfftxServer*: argument checks before any device work; null handles."""
    import ctypes

    import torch

    L = ff.lib()
    h = ctypes.c_void_p()
    assert L.flagfftxServerCreate(None, 1 << 20, 2, 0) == R.INVALID_VALUE
    assert L.flagfftxServerCreate(ctypes.byref(h), 0, 2, 0) == R.INVALID_VALUE
    assert L.flagfftxServerCreate(ctypes.byref(h), 1 << 20, 0, 0) == R.INVALID_VALUE
    assert L.flagfftxServerCreate(ctypes.byref(h), 1 << 20, 2, 7) == R.INVALID_VALUE
    if not torch.cuda.is_available():
        assert L.flagfftxServerCreate(ctypes.byref(h), 1 << 20, 2, 0) == R.INVALID_DEVICE and not h.value
    t = ctypes.c_longlong()
    assert L.flagfftxServerSubmit(None, int(T.C2C), 64, 1, -1, 8, ctypes.byref(t)) == R.INVALID_PLAN
    assert L.flagfftxServerWait(None, 1, None, None) == R.INVALID_PLAN
    assert L.flagfftxServerRelease(None, 1) == R.INVALID_PLAN
    assert L.flagfftxServerGetStats(None, None) == R.INVALID_PLAN
    assert L.flagfftxServerDestroy(None) == R.INVALID_PLAN


def test_null_handles():
    L = ff.lib()
    assert L.flagfftDestroy(None) == R.INVALID_PLAN
    assert L.flagfftSetStream(None, None) == R.INVALID_PLAN
    assert L.flagfftGetPlanDescription(None) is None
    assert L.flagfftExecC2C(None, 8, 8, ff.FORWARD) == R.INVALID_PLAN
    assert L.flagfftExecR2C(None, 8, 8) == R.INVALID_PLAN
    assert L.flagfftxPlanStats(None, -1, None, None) == R.INVALID_PLAN


# ---- planner invariants -------------------------------------------------------
def _describe(n, t=T.C2C, batch=1, rank=1):
    r, s = ff.describe_host(rank, n if isinstance(n, list) else [n], t, batch)
    assert r == R.SUCCESS, (n, r)
    return s


def _core_checks(n, text):
    leaf = re.search(r"Leaf\(n=(\d+), radices=\[([\d,]*)\]", text)
    multi = re.search(r"MultiPassStockham\(n=(\d+), super_radices=\[([\d,]+)\]", text)
    if "Rader(" in text:
        m = re.search(r"Rader\(p=(\d+), root=(\d+), conv_length=(\d+)", text)
        p, g, conv = map(int, m.groups())
        assert conv == p - 1 and all(p % d for d in range(2, int(p ** 0.5) + 1))
        assert len({pow(g, e, p) for e in range(1, p)}) == p - 1
    elif "Bluestein(" in text:
        m = re.search(r"Bluestein\(n=(\d+), conv_length=(\d+)", text)
        nn, conv = map(int, m.groups())
        assert conv >= 2 * nn - 1
        c = conv
        for q in (2, 3, 5, 7, 11, 13):
            while c % q == 0:
                c //= q
        assert c == 1
    if multi and "Bluestein" not in text:
        nn = int(multi.group(1))
        parts = [int(v) for v in multi.group(2).split(",")]
        assert math.prod(parts) == nn and all(2 <= v <= 4096 for v in parts)
    for m in re.finditer(r"(?<!super_)radices=\[([\d,]*)\]", text):
        rad = [int(v) for v in m.group(1).split(",") if v]
        assert all(2 <= v <= 17 or v == 19 for v in rad)
    if leaf and "Rader" not in text and "Bluestein" not in text and not multi:
        rad = [int(v) for v in leaf.group(2).split(",") if v]
        assert math.prod(rad) == int(leaf.group(1))


@pytest.mark.parametrize("n", [1, 2, 3, 16, 17, 23, 64, 67, 81, 97, 243, 256, 361, 512, 997, 1009, 1024, 2048,
                               4096, 4099, 6000, 8192, 10007, 10935, 16384, 15015, 65536, 1 << 20, 1 << 24,
                               2 * 4999, 30011])
def test_planner_invariants(n):
    for t in (T.C2C, T.Z2Z):
        _core_checks(n, _describe(n, t))


@pytest.fixture
def untuned(monkeypatch):
    """The planner's own choices, without rows of the tuned-plan database."""
    monkeypatch.setenv("FLAGFFT_TUNE_DISABLE", "1")


def test_planner_choices(untuned):
    assert "Leaf(n=4096, radices=[16,16,16]" in _describe(4096)
    assert "Rader(p=67" in _describe(67)                      # ctest/test_plan.cpp:47-54
    assert "Rader(p=67" in _describe(67, T.R2C)                # :56-61
    assert "Rader(p=67" in _describe([67, 32], T.C2C, rank=2)  # :142-149
    assert "Bluestein(n=997" in _describe(997)                 # 996 = 4*3*83: not smooth
    assert "Rader(p=1009" in _describe(1009)                   # 1008 = 16*9*7
    assert "RealEvenFused(n=6000" in _describe(6000, T.R2C)
    assert "RealEvenSplitPass(n=32768" in _describe(32768, T.C2R)
    assert "RealPromoted(n=243" in _describe(243, T.D2Z)
    assert "super_radices=[4096,4096]" in _describe(1 << 24)     # two register-resident 64 x 64 passes
    assert "super_radices=[256,256,256]" in _describe(1 << 24, T.Z2Z)  # FP64: no 4096-point instance
    assert "super_radices=[1024,1024]" in _describe(1 << 20)
    assert "TwoDim(n0=4096, n1=4096" in _describe([4096, 4096], T.R2C, rank=2)
    # 4096-point columns: one register-resident pass when its 8-column tiles
    # fill the CUs in whole rounds (4096 columns); the 2049 half-spectrum
    # columns of one real transform take it with the DC and Nyquist columns
    # sharing a unit (2048 units)
    assert "super_radices=[64,64]" not in _describe([4096, 4096], T.C2C, rank=2)
    for t in (T.R2C, T.C2R):
        d = _describe([4096, 4096], t, rank=2)
        assert "dc_nyquist_pair" in d and "super_radices=[64,64]" not in d, d
    # 10^k and 5^k lengths: chains of the 25 / 100 / 125 / 1000-point LDS
    # passes; a Bluestein convolution keeps a power-of-two length (those
    # passes run well below the power-of-two ones)
    assert "Leaf(n=10000, radices=[10,10,10,10]" in _describe(10000, batch=100)  # FP32: 80 KB in LDS
    assert "super_radices=[100,100]" in _describe(10000, T.Z2Z, batch=100)
    assert "super_radices=[1000,1000]" in _describe(1000000, batch=4)
    assert "super_radices=[125,25,25]" in _describe(78125, batch=16)
    assert "conv_length=1048576" in _describe(419017, batch=64)
    assert "conv_length=262144" in _describe(116724, batch=96)
    # 8192-point columns: no single-pass instance with wide row segments
    for t in (T.C2C, T.R2C, T.Z2Z):
        assert "super_radices=[128,64]" in _describe([8192, 8192], t, rank=2)
    assert "dc_nyquist_pair" in _describe([4096, 8192], T.R2C, rank=2)
    assert "dc_nyquist_pair" not in _describe([4096, 4094], T.R2C, rank=2)  # 2048 columns: whole rounds
    assert "dc_nyquist_pair" not in _describe([4096, 4096], T.C2C, rank=2)
    assert "dc_nyquist_pair" not in _describe([2048, 4096], T.R2C, rank=2)  # no 2048-point column instance


def test_tuning_overrides_reach_the_planner(untuned):
    try:
        assert ff.set_tuning(radices=[4, 16, 16, 4]) == R.SUCCESS
        assert "radices=[4,16,16,4]" in _describe(4096)
        assert ff.set_tuning(leaf_max=1024, passes=[64, 64]) == R.SUCCESS
        assert "super_radices=[64,64]" in _describe(4096)
        assert ff.set_tuning(ept=12) == R.INVALID_VALUE
    finally:
        ff.set_tuning()
    assert "Leaf(n=4096, radices=[16,16,16]" in _describe(4096)


def test_tuned_plan_database_round_trip(tmp_path, monkeypatch):
    """A stored row changes the planner's choice for its key only; DISABLE turns lookups off."""
    import sqlite3

    db = tmp_path / "tuned_plans.sqlite"
    monkeypatch.setenv("FLAGFFT_TUNE_DB", str(db))
    monkeypatch.delenv("FLAGFFT_TUNE_DISABLE", raising=False)
    assert ff.tuned_db_path() == str(db)
    ff.set_tuning()
    default = _describe(1 << 16, T.C2C, 4)
    assert "super_radices=[256,256]" in default and "tuned-plan" not in default
    assert ff.tuned_store(1 << 16, T.C2C, 4, 10.0, 100.0, 12.0, passes=[1024, 64]) == R.SUCCESS
    tuned = _describe(1 << 16, T.C2C, 4)
    assert "super_radices=[1024,64]" in tuned and "tuned-plan:" in tuned
    # same batch bucket (2-8) matches; other buckets, types and lengths do not
    assert "super_radices=[1024,64]" in _describe(1 << 16, T.C2C, 8)
    assert "super_radices=[256,256]" in _describe(1 << 16, T.C2C, 64)
    assert "super_radices=[256,256]" in _describe(1 << 16, T.Z2Z, 4) or "tuned-plan" not in _describe(1 << 16, T.Z2Z, 4)
    # newest row wins
    assert ff.tuned_store(1 << 16, T.C2C, 4, 9.0, 110.0, 12.0, passes=[64, 1024]) == R.SUCCESS
    assert "super_radices=[64,1024]" in _describe(1 << 16, T.C2C, 4)
    # explicit thread tuning beats the database
    ff.set_tuning(passes=[512, 128])
    assert "super_radices=[512,128]" in _describe(1 << 16, T.C2C, 4)
    ff.set_tuning()
    monkeypatch.setenv("FLAGFFT_TUNE_DISABLE", "1")
    assert "tuned-plan" not in _describe(1 << 16, T.C2C, 4)
    monkeypatch.delenv("FLAGFFT_TUNE_DISABLE")
    # the table has the reference's key columns
    with sqlite3.connect(db) as c:
        cols = [r[1] for r in c.execute("PRAGMA table_info(tuned_measurements)")]
        rows = c.execute("SELECT fft_length, batch_bucket, dtype, direction, norm, input_layout, status "
                         "FROM tuned_measurements").fetchall()
    for col in ("plan_json", "schema_version", "status", "rank", "device_arch", "fft_length", "batch_bucket", "dtype",
                "direction", "norm", "input_layout", "planner_fingerprint", "codegen_fingerprint",
                "runtime_fingerprint", "measured_at"):
        assert col in cols
    assert rows[0] == (1 << 16, "2-8", "complex64", "forward", "backward", "contiguous", "valid")
    # invalid overrides are refused before anything is written
    assert ff.tuned_store(1 << 16, T.C2C, 4, 1.0, 1.0, 1.0, ept=7) == R.INVALID_VALUE
