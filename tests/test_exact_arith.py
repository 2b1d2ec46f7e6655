"""The division- and sqrt-free forms of hm_exact.h against the IEEE operations
they replace, on the host: tests/exact_arith.cpp includes the same header the
device code does and checks each helper over its whole domain where that is
finite (i / n for n <= 128, every DOM depth code) and densely elsewhere (the
sqrt bound, quotient signs, DOM slices near every integer boundary)."""
from __future__ import annotations

import os
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


@pytest.mark.skipif(shutil.which("g++") is None, reason="no host C++ compiler")
def test_exact_forms_equal_ieee_operations(tmp_path):
    exe = str(tmp_path / "exact_arith")
    subprocess.run(["g++", "-O2", "-std=c++17", "-ffp-contract=off", "-I", os.path.join(ROOT, "bevy_hair_amd", "csrc"),
                    os.path.join(ROOT, "tests", "exact_arith.cpp"), "-o", exe], check=True)
    out = subprocess.run([exe], capture_output=True, text=True, timeout=120)
    assert out.returncode == 0 and out.stdout.strip().endswith("ok"), out.stdout + out.stderr
