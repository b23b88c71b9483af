"""Runs the C++ operator-registry tests (tests/cpp/test_engine_v2.cpp)."""
import os
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
BIN = os.path.join(ROOT, "tests", "cpp", "build", "test_engine_v2")


def _run(*args):
    if not os.path.exists(BIN):
        pytest.fail(f"{BIN} is not built (build() compiles it)")
    p = subprocess.run([BIN, *args], capture_output=True, text=True, timeout=120)
    assert p.returncode == 0, p.stdout + p.stderr


def test_registry_and_capability_queries():
    _run("--cpu")


@pytest.mark.gpu
def test_gpu_operators_through_the_registry():
    _run()
