"""Pins the CPU oracle (oracle/oracle.c) before it is trusted as the checker.

Each oracle function is checked against something independent of its own
formula: published SplitMix64 outputs, closed-form DFTs, the long-double
O(N^2) definition, numpy pocketfft fixtures (tests/golden/golden_fft.npz),
brute force on small integers, and transform identities.  A planted mistake
(wrong sign, dropped term) is shown to fail the same checks.
"""
import os

import numpy as np
import pytest

from oracle import oracle as ora

GOLDEN = np.load(os.path.join(os.path.dirname(__file__), "golden", "golden_fft.npz"))


def rel(a, b):
    return float(np.linalg.norm(np.asarray(a) - np.asarray(b)) / max(np.linalg.norm(b), 1e-300))


# ---- inputs ----------------------------------------------------------------
def test_splitmix64_known_answers():
    """SplitMix64 with seed 0 starts 0xe220a8397b1dcdaf, 0x6e789e6aa1b965f4, 0x06c45d188009454f."""
    x = ora.signed_unit(0, 3)
    bits = [int(round((v + 1.0) / 2.0 * 2.0 ** 53)) for v in x]
    assert bits == [0xE220A8397B1DCDAF >> 11, 0x6E789E6AA1B965F4 >> 11, 0x06C45D188009454F >> 11]


def test_accuracy_seed_layout():
    """ctest/flagfft_test.h:413-417."""
    s = ora.accuracy_seed(0x29, 4096, 256, 0)
    assert s == (0x4654464654455354 ^ (0x29 << 48) ^ (4096 << 16) ^ 256)
    assert ora.accuracy_seed(0x29, 4096, 256, 1) == s ^ 0x9E3779B97F4A7C15


def test_rng_moments():
    u = ora.signed_unit(123, 200000)
    assert u.min() >= -1.0 and u.max() < 1.0 and abs(u.mean()) < 0.01 and abs(u.var() - 1 / 3) < 0.01
    z = ora.normal(321, 200001)
    assert abs(z.mean()) < 0.01 and abs(z.var() - 1.0) < 0.02
    assert abs(np.mean(z ** 4) - 3.0) < 0.1


# ---- definition: naive long-double DFT ---------------------------------------
@pytest.mark.parametrize("n", [1, 2, 3, 7, 16, 23, 97, 360])
def test_dft_closed_forms(n):
    k = np.arange(n)
    for j in {0, min(1, n - 1), n // 2, n - 1}:
        delta = np.zeros(n, complex)
        delta[j] = 1
        expect = np.exp(-2j * np.pi * ((j * k) % n) / n)
        assert rel(ora.dft(delta), expect) < 1e-15
        assert rel(ora.dft(delta, sign=1), np.conj(expect)) < 1e-15
    ones = ora.dft(np.ones(n, complex))
    assert abs(ones[0] - n) < 1e-12 * n and np.max(np.abs(ones[1:]), initial=0) < 1e-12 * n
    for m in {0, min(1, n - 1), n - 1}:
        tone = np.exp(2j * np.pi * ((m * k) % n) / n)
        out = ora.dft(tone)
        expect = np.zeros(n, complex)
        expect[m] = n
        assert np.max(np.abs(out - expect)) < 1e-12 * n


def test_dft_matches_numpy_fixtures():
    for key in GOLDEN.files:
        if not key.startswith("c2c_fwd_") or key.endswith("_seed"):
            continue
        n = int(key.split("_")[-1])
        x = ora.complex_signed_unit(int(GOLDEN[key + "_seed"]), (n,))
        assert rel(ora.dft(x), GOLDEN[key]) < 1e-14, key


# ---- fast FP64 / FP32 FFT (the CPU baseline) ---------------------------------
def test_fast_fft_all_small_lengths():
    for n in range(1, 301):
        x = ora.complex_signed_unit(1000 + n, (2, n))
        ref = ora.dft(x)
        assert rel(ora.fft(x), ref) < 2e-15 * max(1, np.log2(n)), n
        assert rel(ora.fft(x, sign=1), ora.dft(x, sign=1)) < 2e-15 * max(1, np.log2(n)), n


@pytest.mark.parametrize("n", [1024, 4096, 4099, 6000, 8192, 10007, 10935, 3 * 5 * 7 * 11 * 13])
def test_fast_fft_reference_sizes(n):
    x = ora.complex_signed_unit(n, (1, n))
    assert rel(ora.fft(x), ora.dft(x)) < 1e-14


def test_fast_fft_matches_golden():
    for key in GOLDEN.files:
        if key.endswith("_seed"):
            continue
        seed = int(GOLDEN[key + "_seed"])
        kind = key.rsplit("_", 1)[0]
        if kind in ("c2c_fwd", "c2c_inv"):
            n = int(key.split("_")[-1])
            x = ora.complex_signed_unit(seed, (n,))
            got = ora.fft(x, sign=-1 if kind == "c2c_fwd" else 1)
        elif kind == "r2c":
            n = int(key.split("_")[-1])
            got = ora.rfft(ora.signed_unit(seed, n))
        elif kind == "c2r":
            n = int(key.split("_")[-1])
            got = ora.irfft(ora.complex_signed_unit(seed, (n // 2 + 1,)), n)
        elif kind == "fft2":
            n0, n1 = map(int, key.split("_")[-1].split("x"))
            got = ora.fft2(ora.complex_signed_unit(seed, (n0, n1)))
        elif kind == "rfft2":
            n0, n1 = map(int, key.split("_")[-1].split("x"))
            got = ora.rfft2(ora.signed_unit(seed, n0 * n1).reshape(n0, n1))
        else:
            raise AssertionError(key)
        assert rel(got, GOLDEN[key]) < 1e-14, key


def test_fast_fft_f32_error_level():
    for n in [64, 1000, 4096, 10007]:
        x = ora.complex_signed_unit(7 * n, (2, n)).astype(np.complex64)
        e = rel(ora.fft_f32(x), ora.dft(x.astype(np.complex128)))
        assert e < 2e-7 * max(1, np.log2(n)), (n, e)


def test_identities_parseval_roundtrip_linearity():
    n = 6000
    x = ora.complex_normal(5, (3, n))
    y = ora.complex_normal(6, (3, n))
    X = ora.fft(x)
    assert abs(np.sum(np.abs(X) ** 2) / n - np.sum(np.abs(x) ** 2)) < 1e-10 * np.sum(np.abs(x) ** 2)
    assert rel(ora.fft(X, sign=1) / n, x) < 1e-14
    assert rel(ora.fft(2 * x - 3j * y), 2 * X - 3j * ora.fft(y)) < 1e-14
    # shift theorem: x[j - s] <-> X[k] w^{sk}
    s = 17
    k = np.arange(n)
    assert rel(ora.fft(np.roll(x, s, axis=1)), X * np.exp(-2j * np.pi * ((s * k) % n) / n)) < 1e-13


def test_planted_mistakes_are_caught():
    """A wrong sign, a dropped term or a transposed 2D axis must fail the checks used above."""
    n = 997
    key = f"c2c_fwd_{n}"
    x = ora.complex_signed_unit(int(GOLDEN[key + "_seed"]), (n,))
    assert rel(ora.fft(x, sign=1), GOLDEN[key]) > 0.5
    x_dropped = x.copy()
    x_dropped[n // 3] = 0
    assert rel(ora.fft(x_dropped), GOLDEN[key]) > 1e-3
    x2 = ora.complex_signed_unit(int(GOLDEN["fft2_64x32_seed"]), (64, 32))
    wrong = ora.fft(x2)  # rows only, column pass missing
    assert rel(wrong, GOLDEN["fft2_64x32"]) > 0.5


# ---- real transforms -----------------------------------------------------------
@pytest.mark.parametrize("n", [1, 2, 3, 15, 16, 67, 6000])
def test_r2c_is_half_spectrum(n):
    x = ora.signed_unit(n, 3 * n).reshape(3, n)
    full = ora.dft(x.astype(np.complex128))
    assert rel(ora.rfft(x), full[:, : n // 2 + 1]) < 1e-14


@pytest.mark.parametrize("n", [2, 3, 15, 16, 6000])
def test_c2r_drops_dc_and_nyquist_imaginary(n):
    X = ora.complex_signed_unit(n + 1, (2, n // 2 + 1))
    out = ora.irfft(X, n)
    Xh = X.copy()
    Xh[:, 0] = Xh[:, 0].real
    if n % 2 == 0:
        Xh[:, -1] = Xh[:, -1].real
    assert rel(out, ora.irfft(Xh, n)) < 1e-15
    # and equals Re IDFT of the Hermitian extension
    full = np.concatenate([Xh, np.conj(Xh[:, 1:(n + 1) // 2][:, ::-1])], axis=1)
    assert rel(out, ora.dft(full, sign=1).real) < 1e-14
    # roundtrip R2C -> C2R = n x
    x = ora.signed_unit(n + 7, 2 * n).reshape(2, n)
    assert rel(ora.irfft(ora.rfft(x), n), n * x) < 1e-14


def _hermitian_full(X, n):
    """Full n-point spectrum of a Hermitian half spectrum (Im of DC / Nyquist dropped)."""
    Xh = X.copy()
    Xh[..., 0] = Xh[..., 0].real
    if n % 2 == 0:
        Xh[..., -1] = Xh[..., -1].real
    return np.concatenate([Xh, np.conj(Xh[..., 1:(n + 1) // 2][..., ::-1])], axis=-1)


@pytest.mark.parametrize("n", [2, 3, 15, 16, 67, 1000, 4096, 6000, 10007])
def test_fast_real_f32_against_the_long_double_dft(n):
    """FP32 R2C / C2R (the FP32 real CPU-baseline path) against the long-double
    DFT of the same FP32 data, at the FP32 error level."""
    bound = 2e-7 * max(1.0, np.log2(n))
    x = ora.signed_unit(11 * n, 3 * n).reshape(3, n).astype(np.float32)
    X = ora.rfft_f32(x)
    assert X.dtype == np.complex64 and X.shape == (3, n // 2 + 1)
    assert rel(X, ora.dft(x.astype(np.complex128))[:, : n // 2 + 1]) < bound
    H = ora.complex_signed_unit(13 * n, (3, n // 2 + 1)).astype(np.complex64)
    y = ora.irfft_f32(H, n)
    assert y.dtype == np.float32 and y.shape == (3, n)
    assert rel(y, ora.dft(_hermitian_full(H.astype(np.complex128), n), sign=1).real) < bound
    assert rel(ora.irfft_f32(X, n), n * x.astype(np.float64)) < 2 * bound


def test_fast_real_f32_matches_golden():
    """FP32 R2C / C2R against the numpy (pocketfft) fixtures of the FP64 data,
    within the FP32 rounding of the input and the transform."""
    for key in GOLDEN.files:
        if key.endswith("_seed") or not key.startswith(("r2c_", "c2r_")):
            continue
        n = int(key.split("_")[-1])
        seed = int(GOLDEN[key + "_seed"])
        if key.startswith("r2c_"):
            got = ora.rfft_f32(ora.signed_unit(seed, n).astype(np.float32))
        else:
            got = ora.irfft_f32(ora.complex_signed_unit(seed, (n // 2 + 1,)).astype(np.complex64), n)
        assert rel(got, GOLDEN[key]) < 3e-7 * max(1.0, np.log2(n)), key


def test_fast_real_f32_planted_mistakes_are_caught():
    """A conjugated spectrum, a dropped input sample, a Nyquist bin counted
    twice and a reversed output each fail the checks above; C2R ignores a large
    imaginary Nyquist part."""
    n = 6000
    x = ora.signed_unit(99, n).reshape(1, n).astype(np.float32)
    ref = ora.dft(x.astype(np.complex128))[:, : n // 2 + 1]
    X = ora.rfft_f32(x)
    assert rel(X, ref) < 1e-6
    assert rel(np.conj(X), ref) > 0.5
    xd = x.copy()
    xd[0, 1234] = 0
    assert rel(ora.rfft_f32(xd), ref) > 1e-3
    H = ora.complex_signed_unit(7, (1, n // 2 + 1))
    H[0, -1] += 1000j  # a large imaginary Nyquist part that C2R must ignore
    want = ora.dft(_hermitian_full(H, n), sign=1).real
    got = ora.irfft_f32(H.astype(np.complex64), n)
    assert rel(got, want) < 1e-6
    # the Nyquist bin mirrored as if it had a partner (counted twice)
    twice = _hermitian_full(H, n)
    twice[:, n // 2] *= 2
    assert rel(ora.dft(twice, sign=1).real, want) > 1e-3
    assert rel(got[:, ::-1], want) > 0.5


def test_c2r_zero_is_exact_zero():
    """ctest/test_exec_c2r_ct_s.cpp:76-96."""
    out = ora.irfft(np.zeros((1, 129), complex), 256)
    assert np.all(out == 0.0)


def test_2d_separability():
    for n0, n1 in [(23, 16), (16, 23), (8, 12)]:
        x = ora.complex_signed_unit(n0 * n1, (2, n0, n1))
        rows = ora.dft(x)
        both = np.swapaxes(ora.dft(np.ascontiguousarray(np.swapaxes(rows, 1, 2))), 1, 2)
        assert rel(ora.fft2(x), both) < 1e-14
        assert rel(ora.fft2(ora.fft2(x), sign=1) / (n0 * n1), x) < 1e-14
        xr = x.real.copy()
        assert rel(ora.rfft2(xr), np.fft.fft2(xr)[..., : n1 // 2 + 1]) < 1e-14
        assert rel(ora.irfft2(ora.rfft2(xr), n1), n0 * n1 * xr) < 1e-14


def test_3d_compositions():
    """fft3 / rfft3 / irfft3 against numpy's pocketfft (a library routine) and
    the closed form of a shifted delta; a transposed axis order would fail
    the non-cubic shapes."""
    for n0, n1, n2 in [(4, 6, 10), (7, 5, 3), (2, 9, 8)]:
        x = ora.complex_signed_unit(n0 * n1 * n2 + 1, (2, n0, n1, n2))
        assert rel(ora.fft3(x), np.fft.fftn(x, axes=(1, 2, 3))) < 1e-14
        assert rel(ora.fft3(x, sign=1), np.fft.ifftn(x, axes=(1, 2, 3)) * n0 * n1 * n2) < 1e-14
        xr = x.real.copy()
        assert rel(ora.rfft3(xr), np.fft.rfftn(xr, axes=(1, 2, 3))) < 1e-14
        assert rel(ora.irfft3(ora.rfft3(xr), n2), n0 * n1 * n2 * xr) < 1e-14
        d = np.zeros((n0, n1, n2), complex)
        a, b, c = 1, 2 % n1, 3 % n2
        d[a, b, c] = 1
        k0, k1, k2 = np.meshgrid(np.arange(n0), np.arange(n1), np.arange(n2), indexing="ij")
        wave = np.exp(-2j * np.pi * (k0 * a / n0 + k1 * b / n1 + k2 * c / n2))
        assert rel(ora.fft3(d), wave) < 1e-14


# ---- Rader / Bluestein restatements --------------------------------------------
def _brute_primitive_root(p):
    for g in range(2, p):
        if len({pow(g, e, p) for e in range(1, p)}) == p - 1:
            return g
    return 1


def test_primitive_root_and_index_table():
    primes = [q for q in range(3, 400) if all(q % d for d in range(2, int(q ** 0.5) + 1))]
    for p in primes:
        g = ora.primitive_root(p)
        assert g == _brute_primitive_root(p), p
        idx = ora.rader_index_table(p, g)
        assert sorted(idx.tolist()) == list(range(1, p)), p
    assert ora.primitive_root(10007) == 5 and ora.primitive_root(1009) == 11
    assert ora.primitive_root(12) == -1


def test_bluestein_chirp_exact_reduction():
    for n in [23, 997, 10007]:
        c = ora.bluestein_chirp(n, -1)
        k = np.arange(n)
        r = np.array([(int(v) * int(v)) % (2 * n) for v in k], dtype=np.float64)
        assert rel(c, np.exp(-1j * np.pi * r / n)) < 1e-15
        assert rel(ora.bluestein_chirp(n, 1), np.conj(c)) == 0.0


@pytest.mark.parametrize("p", [5, 23, 67, 997, 1009])
def test_rader_restatement(p):
    x = ora.complex_signed_unit(p, (2, p))
    for sign in (-1, 1):
        assert rel(ora.rader_fft(x, sign), ora.dft(x, sign=sign)) < 1e-14


@pytest.mark.parametrize("n,m", [(23, 45), (23, 64), (997, 2000), (997, 2048), (1009, 4096)])
def test_bluestein_restatement(n, m):
    x = ora.complex_signed_unit(n + m, (2, n))
    for sign in (-1, 1):
        assert rel(ora.bluestein_fft(x, m, sign), ora.dft(x, sign=sign)) < 1e-14


# ---- metrics ---------------------------------------------------------------------
def test_rel_l2_hand_cases():
    ref = np.array([[3 + 4j, 0], [1, 0]])
    val = np.array([[3 + 4j, 1], [1, 1]])
    l2, linf = ora.rel_l2(val, ref)
    assert abs(l2 - 1.0) < 1e-15 and abs(linf - 1.0) < 1e-15  # worst row: |1| / |1|
    l2, linf = ora.rel_l2(np.array([[3.0, 4.0]]), np.array([[3.0, 4.0]]))
    assert l2 == 0.0 and linf == 0.0
    assert ora.work_factor(64) == 64 and ora.work_factor(4096) == 3 * 13 + 3
