"""Device joints against the reference World with the CUDA backend's joint
path (bit for bit).

`joints` is written by the reference's own World::step (oracle/ref_harness.cpp,
`fixture joints`) with a backend that runs joints as backend_cuda.cu does
(64 or more joints: joint_solver.h colour by colour in the device substeps):
8 chains of 8 ball-jointed boxes, two of those joints breakable, and one
hinge (motor), prismatic (limits, motor), rigid distance, spring, rope, fixed,
weld, motor, pulley, gear, wheel and cone-twist joint.  Every listed step
compares body states, the joint records (impulses, reactions, broken flags,
order after removals) and contact events.

`joints_trig` holds the types whose solve calls atan2f or acosf (hinge with
limits, cone-twist with swing and twist limits, six-degree-of-freedom limits
and motors, soft weld, wheel steering) beside 64 chain joints.  The device's
atan2f/acosf are not glibc's, so those bodies are held to a tolerance and the
chain bodies, which no trig result reaches, to the bit.  
`substeps_joints` is one advanceSubsteps call as the reference's World::step
makes it (step 61 of the `joints` scene: 7 contacts, 75 joints, 4 substeps),
its arguments and what the CUDA backend's path leaves in them; the per-call
entry point must reproduce it bit for bit.  Needs an MI355X."""
import numpy as np
import pytest

from vxfix import (as_bodies, as_contacts, as_events, as_state, bits_equal, body_state, contact_bytes,
                   default_options, load_case)
from velox_amd.records import JOINT

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


# Bodies 1..72 are the eight 9-body chains; 73..82 the five trig-joint pairs
# (hinge 74, cone-twist 76, six-dof 78, soft weld 80, wheel 82 are dynamic).
_CHAIN_BODIES = slice(0, 73)
_WELD_BODY = 80
# atan2f / acosf of the device library and of glibc differ by a few ulp.  The
# cone-twist body meets its limits and amplifies any such difference: the
# reference run against itself with its initial spin moved by 2e-6 rad/s
# separates by 8e-5 in position and 7e-3 in angular velocity at step 30.
_TOL = {"p": 1e-4, "q": 1e-4, "v": 1e-2, "w": 1e-2}
# The reference's soft weld solve feeds its accumulated impulse back without
# the softness term in the denominator (joint_solver.h:456-469) and spins its
# body up to 120 rad/s by step 10; past that step the trajectory depends on
# every ulp, so the weld body is compared through step 10 only.
_WELD_LAST_STEP = 10


def test_trig_joint_types_track_reference(ctx):
    case = load_case("joints_trig")
    meta = case["meta"]
    opts = default_options(friction_model=meta["frictionModel"], velocity_iterations=meta["velocityIterations"],
                           position_iterations=meta["positionIterations"])
    ctx.load(as_bodies(case["bodies_in"]), case["soup"], gravity=tuple(meta["gravity"]), substeps=meta["substeps"],
             options=opts)
    ctx.set_joints(case["joints_in"])
    done = 0
    for k, target in enumerate(meta["steps"]):
        print(f"[gpu] joints_trig step {target}", flush=True)
        ctx.step(np.float32(meta["dt"]), target - done)
        done = target
        got, want = body_state(ctx.bodies()), as_state(case[f"state_{target}"])
        assert len(got) == len(want) == 83
        assert bits_equal(got[_CHAIN_BODIES], want[_CHAIN_BODIES]), f"step {target}: chain bodies differ"
        bodies = [b for b in range(73, 83) if b != _WELD_BODY or target <= _WELD_LAST_STEP]
        for f, tol in _TOL.items():
            g, w = got[bodies][f], want[bodies][f]
            assert np.all(np.isfinite(got[73:83][f])), f"step {target}: non-finite {f}"
            per_body = np.abs(g - w).max(axis=1)
            err = float(per_body.max())
            print(f"[gpu]   {f}: max |device - reference| = {err:.3g} (body {bodies[int(per_body.argmax())]})",
                  flush=True)
            if err > tol:
                pytest.fail(f"step {target}: {f} off by {err}")
        assert len(ctx.joints()) == meta["joints_at"][k]


def _substeps_case():
    case = load_case("substeps_joints")
    bodies = as_bodies(case["bodies_in"])
    contacts = as_contacts(case["contacts_in"])
    joints = np.ascontiguousarray(case["joints_in"]).view(JOINT).reshape(-1).copy()
    return case, bodies, contacts, joints


def test_advance_substeps_with_joints_matches_reference(ctx):
    case, bodies, contacts, joints = _substeps_case()
    meta = case["meta"]
    assert len(contacts) == meta["contacts"] and len(joints) == meta["joints"] >= 64
    opts = default_options(friction_model=meta["frictionModel"], velocity_iterations=meta["velocityIterations"])
    print(f"[gpu] advance_substeps: {len(bodies)} bodies, {len(contacts)} contacts, {len(joints)} joints", flush=True)
    assert ctx.advance_substeps(bodies, contacts, tuple(meta["gravity"]), np.float32(meta["h"]), meta["substeps"],
                                opts, joints=joints)
    want_b = as_bodies(case["bodies_out"])
    fields = [f for f in bodies.dtype.names if not bits_equal(bodies[f], want_b[f])]
    for f in fields:
        rows = np.nonzero([not bits_equal(x, y) for x, y in zip(bodies[f], want_b[f])])[0]
        print(f"[gpu]   body field {f} differs in {len(rows)} bodies, first {rows[:4].tolist()}: "
              f"{bodies[f][rows[0]]} vs {want_b[f][rows[0]]}", flush=True)
    assert not fields, f"body fields differ: {fields}"
    want_c = as_contacts(case["contacts_out"])
    assert np.array_equal(contact_bytes(contacts), contact_bytes(want_c)), "contacts differ"
    got_j = joints.view(np.uint8).reshape(-1, 420)
    want_j = np.ascontiguousarray(case["joints_out"]).reshape(-1, 420)
    assert np.array_equal(got_j, want_j), f"joint records differ at {_bad(got_j, want_j)}"
    assert ctx.last_velocity_iterations() == 2 * meta["velocityIterations"] * meta["substeps"]


def test_advance_substeps_declines_like_the_cuda_backend(ctx):
    case, bodies, contacts, joints = _substeps_case()
    meta = case["meta"]
    g, h, S = tuple(meta["gravity"]), np.float32(meta["h"]), meta["substeps"]
    # 1..63 joints: the host solves them (backend_cuda.cu:696-698); nothing is touched.
    before = bodies.copy()
    assert not ctx.advance_substeps(bodies, contacts, g, h, S, default_options(), joints=joints[:63].copy())
    assert bits_equal(bodies, before)
    # A joint type past Gear (backend_cuda.cu:699-700).
    odd = joints.copy()
    odd["type"][5] = 13
    assert not ctx.advance_substeps(bodies, contacts, g, h, S, default_options(), joints=odd)
    assert bits_equal(bodies, before)
    # Adaptive is declined after joints between an asleep and an awake body
    # woke both (backend_cuda.cu:703-715).
    dyn = [k for k in range(len(joints)) if bodies["invMass"][joints["a"][k]] > 0
           and bodies["invMass"][joints["b"][k]] > 0]
    k = dyn[0]
    a, b = int(joints["a"][k]), int(joints["b"][k])
    bodies["asleep"][a] = 1
    bodies["sleepTimer"][a] = np.float32(0.75)
    opts = default_options()
    opts.iterationPolicy = 1
    assert not ctx.advance_substeps(bodies, contacts, g, h, S, opts, joints=joints)
    assert bodies["asleep"][a] == 0 and bodies["asleep"][b] == 0
    assert bodies["sleepTimer"][a] == 0 and bodies["sleepTimer"][b] == 0
