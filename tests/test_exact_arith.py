"""The kernels' division-free exact arithmetic (sqrt_rn, the TSDF decode, the
raycast's division by the cell size and the TSDF running average's division
by the weight) equals the IEEE operations it
replaces for every input of its range: an exhaustive CPU check
(tests/cpp/exact_arith_check.c), compiled and run here."""
import os
import shutil
import subprocess

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))


@pytest.mark.skipif(shutil.which("gcc") is None, reason="needs gcc")
def test_exact_arithmetic_exhaustive(tmp_path):
    exe = str(tmp_path / "exact_arith_check")
    src = os.path.join(HERE, "cpp", "exact_arith_check.c")
    subprocess.run(["gcc", "-O2", "-march=native", "-fopenmp", "-ffp-contract=off", "-o", exe, src, "-lm"],
                   check=True)
    r = subprocess.run([exe], capture_output=True, text=True, timeout=600)
    print(r.stdout)
    assert r.returncode == 0, r.stdout + r.stderr
    assert r.stdout.count(" 0 mismatches") == 7
