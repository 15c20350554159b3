"""Runs the C++ drop-in layer tests (tests/cpp/test_dropin.cpp).

The program is built by __graft_entry__.build() (make -C tests/cpp); each
case prints PASS/FAIL and the exit status counts failures.
"""
import os
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
EXE = os.path.join(ROOT, "tests", "cpp", "_build", "test_dropin")


def _run(*args, timeout=300):
    if not os.path.exists(EXE):
        subprocess.run(["make", "-C", os.path.join(ROOT, "tests", "cpp"), "-s"], check=True)
    r = subprocess.run([EXE, *args], capture_output=True, text=True, timeout=timeout)
    print(r.stdout)
    print(r.stderr)
    return r


def test_dropin_parameters():
    r = _run("params")
    assert r.returncode == 0, r.stdout + r.stderr
    assert r.stdout.count("PASS") == 4


@pytest.mark.gpu
def test_dropin_pipeline_on_dataset_1(tsf):
    r = _run("gpu", os.path.join(ROOT, "tests", "golden", "dataset_1"), timeout=600)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "FAIL" not in r.stdout
    assert r.stdout.count("PASS") == 17


@pytest.mark.gpu
def test_dropin_run_ingest_and_output(tsf):
    r = _run("run", os.path.join(ROOT, "tests", "golden", "dataset_1"), timeout=600)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "FAIL" not in r.stdout
    assert r.stdout.count("PASS") == 13


@pytest.mark.gpu
def test_dropin_autograd_function(tsf):
    r = _run("autograd", timeout=300)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "FAIL" not in r.stdout
    assert r.stdout.count("PASS") == 3
