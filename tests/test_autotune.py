"""Autotuner candidate generation (CPU; timing and storing need a GPU)."""
import numpy as np
import pytest

from flagfft_amd import autotune as at
from flagfft_amd import flagfft as ff

T = ff.Type


def _describe(n, t, batch):
    r, s = ff.describe_host(1, [n], t, batch)
    assert r == ff.Result.SUCCESS
    return s


def test_radix_packing_matches_the_planner():
    for n in (16, 256, 1000, 3000, 4096, 10935, 6561, 17, 361, 4352, 5168):
        rad = at._radices(n)
        assert int(np.prod(rad)) == n and all(2 <= r <= 17 or r == 19 for r in rad)
        assert f"radices=[{','.join(map(str, rad))}]" in _describe(n, T.C2C, 8)
    assert at._radices(23) is None and at._radices(17 * 23) is None


@pytest.mark.parametrize("n,k", [(1 << 20, 2), (1 << 24, 2), (10935, 2), (20480, 3)])
def test_splits_multiply_to_n(n, k):
    splits = at._splits(n, 4096, k)
    assert splits
    for s in splits:
        assert len(s) == k and int(np.prod(s)) == n and max(s) <= 4096
        assert s == sorted(s, reverse=True)


def test_candidates_cover_the_free_choices():
    leaf = at.candidates(T.C2C, 4096, 256, _describe(4096, T.C2C, 256))
    assert leaf[0] == {} and {"units": 2} in leaf and any("radices" in c for c in leaf)
    multi = at.candidates(T.C2C, 1 << 20, 16, _describe(1 << 20, T.C2C, 16))
    passes = [c["passes"] for c in multi if "passes" in c]
    assert [1024, 1024] in passes and any(len(p) == 3 for p in passes)
    real = at.candidates(T.R2C, 1 << 21, 4, _describe(1 << 21, T.R2C, 4))
    assert all(int(np.prod(c["passes"])) == 1 << 20 for c in real if "passes" in c)
    # every candidate is accepted by the planner
    for c in leaf + multi:
        ff.set_tuning(**c) if c else ff.set_tuning()
        try:
            assert ff.describe_host(1, [4096 if c in leaf else 1 << 20], T.C2C, 16)[0] == ff.Result.SUCCESS
        finally:
            ff.set_tuning()
