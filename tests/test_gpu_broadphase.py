"""The broad phase's cell choice: when a few bodies are far larger than the
rest (here fast bodies whose swept boxes are metres long, as the spinning
bodies of the late debris steps are), the grid cell follows the rest and those
bodies take the oversize path.  The pair set must not change: it is compared
with the oracle's exact-AABB pairs, and a short trajectory with the oracle's
replay, bit for bit.  Needs an MI355X."""
import ctypes as C

import numpy as np
import pytest

from vxfix import OracleWorld, as_bodies, bits_equal, load_case, oracle

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ctx():
    from velox_amd.device import Context
    c = Context()
    yield c
    c.close()


def _fast_pile():
    """The npile_40 bodies with five dynamic bodies thrown at 300 m/s (swept
    boxes of about 5 m against bodies of at most 1 m)."""
    case = load_case("npile_40")
    bodies = as_bodies(case["bodies_in"]).copy()
    dyn = np.flatnonzero(bodies["motionType"] == 2)
    fast = dyn[[3, 41, 77, 120, 160]]
    bodies["velocity"][fast] = np.array([300.0, -40.0, 120.0], np.float32)
    return case, bodies, fast


def _canon(pairs):
    a, b = pairs >> np.uint64(32), pairs & np.uint64(0xffffffff)
    return np.sort((np.minimum(a, b) << np.uint64(32)) | np.maximum(a, b))


def test_oversize_fast_bodies_keep_the_pair_set(ctx):
    case, bodies, fast = _fast_pile()
    dt = np.float32(case["meta"]["dt"])
    print(f"[gpu] broad phase with {len(fast)} fast bodies among {len(bodies)}", flush=True)
    got = ctx.probe_pairs(bodies, case["soup"], dt)
    lib = oracle()
    view = case["soup"].view()
    want = np.zeros(64 * len(bodies), np.uint64)
    n = lib.vxo_broadphase(bodies.ctypes.data_as(C.c_void_p), len(bodies), C.byref(view), dt,
                           want.ctypes.data_as(C.c_void_p), len(want))
    assert n < len(want)
    assert np.array_equal(_canon(got), _canon(want[:n]))
    # The fast bodies do pair across many cells' worth of space.
    fastSet = set(int(i) for i in fast)
    lo, hi = got >> np.uint64(32), got & np.uint64(0xffffffff)
    assert sum(1 for x, y in zip(lo, hi) if int(x) in fastSet or int(y) in fastSet) > 0


def test_oversize_fast_bodies_step_like_the_oracle(ctx):
    case, bodies, fast = _fast_pile()
    meta = case["meta"]
    dt = np.float32(meta["dt"])
    ctx.load(bodies, case["soup"], gravity=tuple(meta["gravity"]), substeps=meta["substeps"])
    ref = OracleWorld(bodies, case["soup"], dt=dt, substeps=meta["substeps"], gravity=tuple(meta["gravity"]))
    for t in range(3):
        ctx.step(dt)
        ref.step()
        ctx.sync()
        assert bits_equal(ctx.bodies(), ref.bodies), f"step {t + 1} differs from the oracle"
    st = ctx.stats()
    # The fast bodies (and the largest few others) went through the oversize path.
    assert st["oversize"] > len(fast), st
