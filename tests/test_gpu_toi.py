"""The multi-TOI scheduler (processHighToi) on the device against the
reference World (bit for bit).

`hightoi` is written by the reference's own World::step (oracle/ref_harness.cpp,
`fixture hightoi`): 24 MotionQuality::High spheres fired at 60-240 m/s into a
wall of static boxes, two pairs of High boxes fired at each other, and Medium
boxes elsewhere.  Each listed step compares body states, contact and event
output, and the number of multi-TOI events the step processed.  Needs an
MI355X."""
import numpy as np
import pytest

from vxfix import as_bodies, as_events, as_state, bits_equal, body_state, default_options, load_case

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ctx():
    from velox_amd.device import Context
    c = Context()
    yield c
    c.close()


def _bad(got, want):
    a8 = np.ascontiguousarray(got).view(np.uint8).reshape(len(got), -1)
    b8 = np.ascontiguousarray(want).view(np.uint8).reshape(len(want), -1)
    return np.nonzero((a8 != b8).any(axis=1))[0].tolist()[:8]


def test_high_quality_bodies_match_reference(ctx):
    case = load_case("hightoi")
    meta = case["meta"]
    opts = default_options(friction_model=meta["frictionModel"], velocity_iterations=meta["velocityIterations"],
                           position_iterations=meta["positionIterations"])
    ctx.load(as_bodies(case["bodies_in"]), case["soup"], gravity=tuple(meta["gravity"]), substeps=meta["substeps"],
             options=opts)
    done = 0
    for k, target in enumerate(meta["steps"]):
        print(f"[gpu] hightoi step {target}", flush=True)
        ctx.step(np.float32(meta["dt"]), target - done)
        done = target
        st = ctx.stats()
        got, want = body_state(ctx.bodies()), as_state(case[f"state_{target}"])
        assert bits_equal(got, want), f"step {target}: states differ at {_bad(got, want)}"
        assert st["multiToiEvents"] == meta["toi_events"][k], f"step {target}: {st['multiToiEvents']} TOI events"
        assert st["contacts"] == meta["contacts_at"][k]
        want_e = as_events(case[f"events_{target}"])
        assert bits_equal(ctx.events(), want_e), f"step {target}: events"


def test_scheduler_off_leaves_bodies_to_the_net(ctx):
    """With enableSubstepSplitting off no event is processed."""
    case = load_case("hightoi")
    meta = case["meta"]
    ctx.load(as_bodies(case["bodies_in"]), case["soup"], gravity=tuple(meta["gravity"]), substeps=meta["substeps"],
             options=default_options())
    ctx.set_multi_toi(enableSubstepSplitting=0)
    ctx.step(np.float32(meta["dt"]), 8)
    assert ctx.stats()["multiToiEvents"] == 0
    ctx.set_multi_toi()
