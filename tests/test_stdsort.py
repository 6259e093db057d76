"""The device's restatement of libstdc++'s std::sort (vx_stdsort.hpp, used to
order the warm-start cache exactly as World::step's std::sort does) against
the host libstdc++ itself: same element order among equal keys."""
import os
import shutil
import subprocess

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))


@pytest.mark.skipif(shutil.which("g++") is None, reason="needs g++")
def test_stdsort_restatement_matches_libstdcxx(tmp_path):
    exe = str(tmp_path / "stdsort_check")
    subprocess.run(["g++", "-O2", "-std=c++17", os.path.join(HERE, "stdsort_check.cpp"), "-o", exe], check=True)
    out = subprocess.run([exe], check=True, capture_output=True, text=True).stdout
    assert "mismatches 0" in out, out
