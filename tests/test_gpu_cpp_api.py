"""Runs tests/cpp/test_cuckoo_filter (the reference's behavioural tests written
against the C++ template drop-in and its rocThrust overloads) on the GPU."""
import os
import subprocess

import pytest

pytestmark = pytest.mark.gpu

EXE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "cpp", "test_cuckoo_filter")


def test_cpp_template_api_behaviour(device):
    assert os.path.exists(EXE), "tests/cpp/test_cuckoo_filter is not built (python cuckoo-filter_amd/build.py)"
    r = subprocess.run([EXE], capture_output=True, text=True, timeout=240)
    print(r.stdout)
    assert r.returncode == 0, r.stdout + r.stderr
    assert r.stdout.count("[  OK  ]") == 15
