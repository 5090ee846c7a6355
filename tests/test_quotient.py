"""Exhaustive proof, over all 2^32 binary32 inputs, that the tiled
projection kernel's quotient paths equal IEEE division div / s bit for bit
(tests/csrc/quotient_check.c)."""
import shutil
import subprocess
from pathlib import Path

import pytest

SRC = Path(__file__).resolve().parent / "csrc" / "quotient_check.c"


@pytest.mark.skipif(shutil.which("gcc") is None, reason="needs gcc")
def test_quotient_paths_are_exact(tmp_path):
    exe = tmp_path / "quotient_check"
    subprocess.run(["gcc", "-O2", "-mfma", "-ffp-contract=off", "-fopenmp", "-o", str(exe),
                    str(SRC), "-lm"], check=True)
    res = subprocess.run([str(exe)], capture_output=True, text=True, timeout=600)
    lines = dict(line.rsplit(" ", 1) for line in res.stdout.strip().splitlines())
    assert lines == {"mismatches s=1": "0", "mismatches s=2": "0", "mismatches s=4": "0",
                     "mismatches s=3": "0"}, res.stdout
    assert res.returncode == 0
