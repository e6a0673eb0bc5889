"""Host checks of sa_math.h identities the walk kernel relies on (built with
g++ from the header the kernels include)."""
import os
import subprocess

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))


def test_walk_math_identities(tmp_path):
    exe = str(tmp_path / "math_check")
    src = os.path.join(HERE, "native", "math_check.cpp")
    r = subprocess.run(["g++", "-O2", "-std=c++17", "-ffp-contract=off", src, "-o", exe],
                       capture_output=True, text=True)
    if r.returncode != 0:
        pytest.fail(r.stderr[-2000:])
    r = subprocess.run([exe], capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stdout + r.stderr
