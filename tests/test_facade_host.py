"""The C++ API's host-side logic without a GPU (tests/cpp/facade_host_check.cpp): virtual
sites in OpenMM's shape (ownership, accessors, errors), the LJ combining rule, and the
integrator's reference defaults (LangevinIntegratorSDM.cpp:48-85)."""
import os
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PROG = os.path.join(ROOT, "tests", "cpp", "facade_host_check")


def test_facade_host_logic():
    if not os.path.exists(PROG):
        pytest.fail("tests/cpp/facade_host_check was not built (__graft_entry__.build())")
    out = subprocess.run([PROG], capture_output=True, text=True, timeout=60)
    assert out.returncode == 0, out.stdout + out.stderr
    assert out.stdout.strip().splitlines()[-1] == "ok"
