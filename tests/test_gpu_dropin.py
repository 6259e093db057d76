"""The drop-in on the reference side: the strict reference World (built from
the reference sources by oracle/ref.mk) steps with the HipBackend adapter of
oracle/ref_hip.cpp installed, every device-side Backend call going to
libvxhip's per-call entry points (integrate, findContacts, advanceSubsteps,
fetchImpulses; world.cpp:4014, 4035, 4201-4211, 4360), and its trajectory is
compared bit for bit with the golden states the reference World wrote with
the device backend's semantics.  Needs an MI355X.
"""
import os
import subprocess

import numpy as np
import pytest

from vxfix import ROOT, as_state, bits_equal, load_case

pytestmark = pytest.mark.gpu

EXE = os.path.join(ROOT, "oracle", "_ref", "ref_hip")


@pytest.mark.parametrize("name", ["lattice64", "pile"])
def test_reference_world_through_hip_backend(name, tmp_path):
    if not os.path.exists(EXE):
        pytest.skip("oracle/_ref/ref_hip is built only where the reference sources are present")
    print(f"[gpu] reference World + HipBackend: {name}", flush=True)
    r = subprocess.run([EXE, str(tmp_path), name], capture_output=True, text=True, timeout=240)
    print(r.stdout, r.stderr, flush=True)
    assert r.returncode == 0
    assert "device substep calls" in r.stdout
    case = load_case(name)
    for target in case["meta"]["steps"]:
        got = np.load(os.path.join(str(tmp_path), f"state_{target}.npy"), allow_pickle=False)
        want = case[f"state_{target}"]
        assert bits_equal(as_state(got), as_state(want)), f"{name} step {target}"
