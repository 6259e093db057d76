"""The C-ABI library loads and exports every entry point include/vxhip.h
declares (no device calls: runs without a GPU)."""
import ctypes as C
import os
import re

import pytest

from velox_amd import device

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols():
    text = open(os.path.join(ROOT, "include", "vxhip.h")).read()
    return sorted(set(re.findall(r"^\s*(?:int|void|uint32_t|const\s+\w+\s*\*)\s*(vxhip_\w+)\s*\(", text, re.M)))


def test_header_declares_expected_surface():
    names = declared_symbols()
    assert len(names) >= 20
    assert set(names) == set(device.SIGNATURES), set(names) ^ set(device.SIGNATURES)


@pytest.mark.skipif(not os.path.exists(device.LIBRARY), reason="libvxhip.so not built")
def test_library_exports_every_declared_symbol():
    lib = C.CDLL(device.LIBRARY)
    missing = [n for n in declared_symbols() if not hasattr(lib, n)]
    assert not missing, missing


@pytest.mark.skipif(not os.path.exists(device.LIBRARY), reason="libvxhip.so not built")
def test_last_error_is_callable_without_device():
    lib = device.library()
    assert isinstance(lib.vxhip_last_error(), bytes)


def test_record_layouts_match_header():
    from velox_amd.records import BODY, CONTACT
    assert BODY.itemsize == 320 and CONTACT.itemsize == 128
    assert BODY.fields["shape"][1] == 232 and BODY.fields["compoundCount"][1] == 296
    assert CONTACT.fields["spinningImpulse"][1] == 116


def test_sleep_records_match_header_and_reference_defaults():
    """vx_sleep_config / vx_sleep_stats byte layouts, and the binding's defaults
    equal SleepConfig's (sleep.h:98-160)."""
    assert C.sizeof(device.SleepConfig) == 36 and C.sizeof(device.SleepStats) == 56
    assert device.SleepStats.totalSleepTransitions.offset == 32
    cfg = device.SleepConfig()
    got = [cfg.linearVelocityThreshold, cfg.angularVelocityThreshold, cfg.accelerationThreshold, cfg.timeToSleep,
           cfg.timeToDrowsy, cfg.drowsySimulationRate, cfg.contactStabilityThreshold]
    assert got == [C.c_float(x).value for x in (0.05, 0.05, 0.01, 0.5, 0.25, 0.25, 0.02)]
    assert (cfg.contactStabilityFrames, cfg.enableGradualSleep, cfg.enableContactStability) == (10, 1, 1)
