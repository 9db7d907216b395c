"""The engine executable (dummest-amd): option handling that needs no GPU.
The end-to-end record parity against the oracle is in test_gpu_cli.py."""
import os
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CLI = os.path.join(ROOT, "dummest_amd", "dummest-amd")


def run(*args, stdin=None):
    return subprocess.run([CLI, *args], input=stdin, capture_output=True, text=True, timeout=60)


def test_help_lists_reference_options():
    r = run("-h")
    assert r.returncode == 0
    for opt in ("--evalue", "--total-length", "--trials", "--length", "--border", "--batch", "--scores-file",
                "--warmup", "--barithmetic", "--bgeometric", "--bmedian", "--insert1", "--insert2", "--delete1",
                "--delete2", "--stop-codon-prob", "--bg-stop-codon-prob", "--max", "--filter-evalue"):
        assert opt in r.stdout, opt


@pytest.mark.parametrize("args", [["-e", "-1"], ["-t", "0"], ["--insert1", "1.5"], ["-m", "7"], ["-m", "-1"],
                                  ["-s", "3"], ["--tantan-threshold", "1.5"], ["--tantan-threshold", "-0.1"]])
def test_bad_option_values_fail(args, tmp_path):
    r = run(*args, str(tmp_path / "x.hmm"))
    assert r.returncode == 1
    assert "usage" in r.stderr or "bad option" in r.stderr.lower() or "invalid" in r.stderr.lower()


@pytest.mark.parametrize("opt", ["-d", "-D"])
def test_gaussian_filter_options(opt, tmp_path):
    # standard deviations outside [2, 1000] are bad options (dummest.cc:185-195);
    # a valid one passes option checking (the run then stops at the missing file)
    for bad in ("1.9", "1001"):
        r = run(opt, bad, str(tmp_path / "x.hmm"))
        assert r.returncode == 1 and "can't open" not in r.stderr
    r = run(opt, "5", str(tmp_path / "x.hmm"))
    assert r.returncode == 1 and "can't open" in r.stderr


@pytest.mark.parametrize("mask", ["0", "1", "2", "3"])
def test_masking_options_accepted(mask, tmp_path):
    # every -m value passes option checking (the run then stops at the missing file)
    r = run("-m", mask, "--tantan-threshold", "0.4", str(tmp_path / "x.hmm"))
    assert r.returncode == 1 and "can't open" in r.stderr


def test_usage_without_arguments():
    r = run()
    assert r.returncode == 1 and "usage" in r.stderr
