"""Device joints against the reference World with the CUDA backend's joint
path (bit for bit).

`joints` is written by the reference's own World::step (oracle/ref_harness.cpp,
`fixture joints`) with a backend that runs joints as backend_cuda.cu does
(64 or more joints: joint_solver.h colour by colour in the device substeps):
8 chains of 8 ball-jointed boxes, two of those joints breakable, and one
hinge (motor), prismatic (limits, motor), rigid distance, spring, rope, fixed,
weld, motor, pulley, gear, wheel and cone-twist joint.  Every listed step
compares body states, the joint records (impulses, reactions, broken flags,
order after removals) and contact events.  Needs an MI355X."""
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


def test_joints_match_reference(ctx):
    case = load_case("joints")
    meta = case["meta"]
    opts = default_options(friction_model=meta["frictionModel"], velocity_iterations=meta["velocityIterations"],
                           position_iterations=meta["positionIterations"])
    ctx.load(as_bodies(case["bodies_in"]), case["soup"], gravity=tuple(meta["gravity"]), substeps=meta["substeps"],
             options=opts)
    ctx.set_joints(case["joints_in"])
    done = 0
    for k, target in enumerate(meta["steps"]):
        print(f"[gpu] joints step {target}", flush=True)
        ctx.step(np.float32(meta["dt"]), target - done)
        done = target
        got, want = body_state(ctx.bodies()), as_state(case[f"state_{target}"])
        assert bits_equal(got, want), f"step {target}: states differ at {_bad(got, want)}"
        got_j = np.ascontiguousarray(ctx.joints()).view(np.uint8).reshape(-1, 420)
        want_j = np.ascontiguousarray(case[f"joints_{target}"]).reshape(-1, 420)
        assert len(got_j) == meta["joints_at"][k] == len(want_j), f"step {target}: joint count"
        assert np.array_equal(got_j, want_j), f"step {target}: joint records differ at {_bad(got_j, want_j)}"
        assert bits_equal(ctx.events(), as_events(case[f"events_{target}"])), f"step {target}: events"


def test_fewer_than_64_joints_stay_on_the_host(ctx):
    from velox_amd.device import VxhipError
    case = load_case("joints")
    meta = case["meta"]
    ctx.load(as_bodies(case["bodies_in"]), case["soup"], gravity=tuple(meta["gravity"]), substeps=meta["substeps"],
             options=default_options())
    with pytest.raises(VxhipError):
        ctx.set_joints(case["joints_in"][:10])
