"""CPU test of the cooperative-panel ledger that keeps concurrent FP64 factorizations' spinning
panel workgroups within the device's CU count (algebra_linear_amd/csrc/coop_ledger.hpp)."""
import os
import shutil
import subprocess

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))


@pytest.mark.skipif(shutil.which("g++") is None, reason="needs g++")
def test_coop_ledger_accounting(tmp_path):
    exe = tmp_path / "coop_ledger_test"
    subprocess.run(["g++", "-std=c++17", "-O1", "-pthread", "-Wall", "-Werror",
                    os.path.join(HERE, "cpp", "coop_ledger_test.cpp"), "-o", str(exe)], check=True)
    out = subprocess.run([str(exe)], capture_output=True, text=True, timeout=60)
    assert out.returncode == 0, out.stderr
    assert out.stdout.strip() == "ok"
