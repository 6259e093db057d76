"""The differentiable form's hand-written adjoints (velox_amd/csrc/
vxd_adjoint.hpp: contact row with its pose-dependent prefix, warm start,
transform advance, renormalisation, positional correction) against central
finite differences of the same maps in double precision, built with
AddressSanitizer and UndefinedBehaviorSanitizer."""
import os
import shutil
import subprocess

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))


@pytest.mark.skipif(shutil.which("g++") is None, reason="needs g++")
def test_adjoints_match_central_differences(tmp_path):
    exe = str(tmp_path / "adjoint_check")
    subprocess.run(["g++", "-O1", "-g", "-std=c++17", "-fsanitize=address,undefined", "-fno-sanitize-recover=all",
                    os.path.join(HERE, "adjoint_check.cpp"), "-o", exe], check=True)
    out = subprocess.run([exe], check=True, capture_output=True, text=True).stdout
    assert "adjoints ok" in out, out
