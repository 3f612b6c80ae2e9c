"""The reference-shaped C++ API (include/ThunderEgg/) on the GPU: tests/cpp/api_test.cpp
builds a forest, a BiQuadratic / DFT V-cycle with GMG::CycleBuilder and solves with
Iterative::BiCGStab, checking the fused device cycle against the generic virtual-call
cycle, mixed precision, the reference's log format and error behaviour;
tests/cpp/patch_api_test.cpp drives the per-patch plug-in points (a user-written
PatchOperator and MPIGhostFiller, the patch solvers around them, a cycle of user
operators, addDrichletBCToRHS, the views) against the built-in device classes."""
import subprocess

import pytest

from test_host import _build_api_test

pytestmark = pytest.mark.gpu


def test_cpp_api_on_device(tmp_path):
    exe = _build_api_test(tmp_path)
    print("case: cpp api_test", flush=True)
    r = subprocess.run([exe], capture_output=True, text=True, timeout=100)
    print(r.stdout)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "api_test: ok" in r.stdout


def test_cpp_patch_plugins_on_device(tmp_path):
    exe = _build_api_test(tmp_path, "patch_api_test")
    print("case: cpp patch_api_test", flush=True)
    r = subprocess.run([exe], capture_output=True, text=True, timeout=100)
    print(r.stdout)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "patch_api_test: ok" in r.stdout
