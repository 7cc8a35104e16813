"""The bench command line (flagfft_amd.cli), after src/cli_tools/flagfft-cli/main.cpp.

CPU: argument validation and exit codes (src/cli_tools/common/cli_utils.hpp:14-17).
GPU: JSON report shape (src/cli_tools/bench/report.cpp:42-67), agreement with
the torch.fft comparator for every API, placement and rank (rank 3 through
flagfftxPlan3d), library errors mapped to the runtime-error exit code.
"""
import json

import pytest

from flagfft_amd import cli


@pytest.mark.parametrize("argv,msg", [
    (["bench", "--rank", "1"], "--shape is required"),
    (["bench", "--shape", "64"], "--rank is required"),
    (["bench", "--rank", "4", "--shape", "8"], "invalid --rank"),
    (["bench", "--rank", "1", "--shape", "4096", "--api", "c2r"], "only supports inverse"),
    (["bench", "--rank", "1", "--shape", "4096", "--api", "r2c", "--direction", "inverse"], "only supports forward"),
    (["bench", "--rank", "1", "--shape", "4x4"], "dimension count"),
    (["bench", "--rank", "2", "--shape", "64x64", "--api", "r2c", "--batch", "2"], "--batch is only supported"),
    (["bench", "--rank", "1", "--shape", "0"], "must be positive"),
    (["bench", "--rank", "1", "--shape", "12a"], "invalid --shape"),
    (["bench", "--rank", "1", "--shape", "64", "--iters", "0"], "--iters must be positive"),
    (["bench", "--rank", "1", "--shape", "64", "--placement", "sideways"], "unknown placement"),
    (["bench", "--rank", "1", "--shape", "64", "--api", "q2q"], "unknown FFT API"),
    (["bench", "--rank", "1", "--shape", "64", "--bogus"], "unknown argument"),
    (["frob"], "unknown command"),
])
def test_bad_arguments_exit_failed(argv, msg, capsys):
    assert cli.main(argv) == cli.EXIT_FAILED
    assert msg in capsys.readouterr().err


def test_parse_multiple_shapes():
    a = cli.parse_args(["bench", "--rank", "2", "--shape", "23x42,128X64", "--api", "z2z", "--batch", "3",
                        "--direction", "inv", "--placement", "ip", "--warmup", "0", "--iters", "5", "--json"])
    assert a["shapes"] == [[23, 42], [128, 64]]
    assert (a["api"], a["batch"], a["direction"], a["placement"]) == ("z2z", 3, "inverse", "in-place")
    assert (a["warmup"], a["iters"], a["json"]) == (0, 5, True)


def test_help_exits_passed(capsys):
    assert cli.main(["--help"]) == cli.EXIT_PASSED
    assert "Usage:" in capsys.readouterr().out


def test_no_device_is_skipped(capsys):
    import torch

    if torch.cuda.is_available():
        pytest.skip("a GPU is present")
    assert cli.main(["bench", "--rank", "1", "--shape", "64"]) == cli.EXIT_SKIPPED
    assert "no HIP device" in capsys.readouterr().err


GPU_CASES = [
    (["--rank", "1", "--shape", "4096,10007", "--batch", "4"], "c2c"),
    (["--rank", "1", "--shape", "1000", "--batch", "3", "--api", "z2z", "--direction", "inverse"], "z2z"),
    (["--rank", "1", "--shape", "6000,243", "--batch", "2", "--api", "r2c"], "r2c"),
    (["--rank", "1", "--shape", "6000", "--batch", "2", "--api", "z2d", "--direction", "inverse"], "z2d"),
    (["--rank", "1", "--shape", "8192", "--batch", "2", "--placement", "in-place"], "c2c"),
    (["--rank", "1", "--shape", "6000,67", "--batch", "3", "--api", "r2c", "--placement", "in-place"], "r2c"),
    (["--rank", "1", "--shape", "256", "--batch", "3", "--api", "c2r", "--direction", "inverse",
      "--placement", "in-place"], "c2r"),
    (["--rank", "2", "--shape", "64x48,23x16", "--batch", "2"], "c2c"),
    (["--rank", "2", "--shape", "128x256", "--api", "d2z"], "d2z"),
    (["--rank", "2", "--shape", "256x130", "--api", "r2c", "--placement", "in-place"], "r2c"),
    (["--rank", "2", "--shape", "64x66", "--api", "c2r", "--direction", "inverse", "--placement", "in-place"], "c2r"),
    (["--rank", "3", "--shape", "16x12x10,9x8x7", "--batch", "2"], "c2c"),
    (["--rank", "3", "--shape", "32x16x20", "--api", "d2z"], "d2z"),
    (["--rank", "3", "--shape", "12x10x9", "--api", "c2r", "--direction", "inverse", "--placement", "in-place"],
     "c2r"),
]


@pytest.mark.gpu
@pytest.mark.parametrize("args,api", GPU_CASES, ids=[f"{i}-{c[1]}" for i, c in enumerate(GPU_CASES)])
def test_bench_json_matches_comparator(gpu, args, api, capsys):
    rc = cli.main(["bench"] + args + ["--warmup", "1", "--iters", "4", "--json", "--check", "--print-path"])
    out = capsys.readouterr()
    assert rc == cli.EXIT_PASSED, out.err
    rep = json.loads(out.out)
    assert rep["status"] == "passed" and rep["command"] == "bench"
    for case in rep["cases"]:
        assert case["api"] == api
        t = case["timing"]
        assert set(t) == {"flagfft_median_ms", "flagfft_p90_ms", "ref_median_ms", "ref_p90_ms", "speedup", "warmup",
                          "iters"}
        assert 0 < t["flagfft_median_ms"] <= t["flagfft_p90_ms"]
        assert t["ref_median_ms"] > 0 and t["speedup"] > 0
        assert case["bytes"]["algorithmic"] > 0 and case["bytes"]["launches"] >= 1
        assert "FlagFFT Plan" in case["plan_description"]
        bound = 1e-12 if api in ("z2z", "d2z", "z2d") else 2e-5
        assert case["rel_l2_vs_ref"] < bound, case


@pytest.mark.gpu
def test_library_errors_are_runtime_errors(gpu, capsys, monkeypatch):
    """A plan the library refuses ends with exit code 2 (cli_utils.cpp check_flagfft)."""
    from flagfft_amd import flagfft as ff

    def refuse(*args, **kw):
        raise ff.FlagFFTError("flagfftPlan1d", int(ff.Result.NOT_SUPPORTED))

    monkeypatch.setattr(ff.Plan, "one_d", refuse)
    assert cli.main(["bench", "--rank", "1", "--shape", "64", "--no-ref"]) == cli.EXIT_RUNTIME_ERROR
    assert "NOT_SUPPORTED" in capsys.readouterr().err


@pytest.mark.gpu
def test_table_output(gpu, capsys):
    assert cli.main(["bench", "--rank", "1", "--shape", "1024", "--batch", "8", "--warmup", "1", "--iters", "3"]) == 0
    lines = capsys.readouterr().out.splitlines()
    assert lines[0].split()[:3] == ["shape", "api", "batch"] and lines[1].split()[:3] == ["1024", "c2c", "8"]
