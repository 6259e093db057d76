"""The device's restatements of libstdc++'s std::sort (vx_stdsort.hpp, and the
level-by-level form of vx_sort.hip that orders the warm-start cache exactly
as World::step's std::sort does) against the host libstdc++ itself: same
element order among equal keys.  Built with AddressSanitizer and
UndefinedBehaviorSanitizer, so an index that leaves its range fails too."""
import os
import shutil
import subprocess

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))


@pytest.mark.skipif(shutil.which("g++") is None, reason="needs g++")
def test_stdsort_restatement_matches_libstdcxx(tmp_path):
    exe = str(tmp_path / "stdsort_check")
    subprocess.run(["g++", "-O1", "-g", "-std=c++17", "-fsanitize=address,undefined", "-fno-sanitize-recover=all",
                    os.path.join(HERE, "stdsort_check.cpp"), "-o", exe], check=True)
    out = subprocess.run([exe], check=True, capture_output=True, text=True).stdout
    assert "mismatches 0" in out, out
