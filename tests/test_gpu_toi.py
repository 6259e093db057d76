"""The multi-TOI scheduler (processHighToi) on the device against the
reference World (bit for bit).

`hightoi` is written by the reference's own World::step (oracle/ref_harness.cpp,
`fixture hightoi`): 24 MotionQuality::High spheres fired at 60-240 m/s into a
wall of static boxes, two pairs of High boxes fired at each other, and Medium
boxes elsewhere.  Each listed step compares body states, contact and event
output, and the number of multi-TOI events the step processed.

`toiquery_<step>` holds World::queryMultiToi for every dynamic body of the same
scene at steps 5, 12 and 20, as the reference returns it (sorted by time of
impact, capped at maxToiEventsPerBody); the device query must give the same
hits bit for bit.  Needs an MI355X."""
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


_HIT_ROW = np.dtype([("query", "<u4"), ("toi", "<f4"), ("body", "<u4"), ("point", "<f4", 3), ("normal", "<f4", 3),
                     ("fraction", "<f4")])


@pytest.mark.parametrize("step", [5, 12, 20])
def test_query_multi_toi_matches_reference(ctx, step):
    case = load_case(f"toiquery_{step}")
    dt = np.float32(case["meta"]["dt"])
    ctx.load(as_bodies(case["bodies_in"]), case["soup"], options=default_options())
    ctx.set_multi_toi()  # the reference defaults (velocity floor 0.1, bias 1e-3)
    rows = np.ascontiguousarray(case["hits"]).view(_HIT_ROW).reshape(-1)
    cap = 4  # min(maxToiEventsPerBody, maxTotalEventsPerStep) at the defaults
    seen = 0
    for q in case["queries"].tolist():
        print(f"[gpu] toiquery step {step} body {q}", flush=True)
        got, total = ctx.query_multi_toi(q, dt)
        want = rows[rows["query"] == q]
        assert len(want) == min(total, cap), f"body {q}: {total} hits, reference {len(want)}"
        for f in ("toi", "body", "point", "normal", "fraction"):
            assert bits_equal(got[f][:len(want)], want[f]), f"body {q}: {f} differs"
        seen += len(want)
    assert seen == case["meta"]["hits"] > 0


def test_query_multi_toi_rejects_bad_arguments(ctx):
    from velox_amd.device import VxhipError
    case = load_case("toiquery_5")
    ctx.load(as_bodies(case["bodies_in"]), case["soup"], options=default_options())
    with pytest.raises(VxhipError):
        ctx.query_multi_toi(0, 0.0)
    with pytest.raises(VxhipError):
        ctx.query_multi_toi(0, float("nan"))
    with pytest.raises(VxhipError):
        ctx.query_multi_toi(len(case["bodies_in"]), 1.0 / 60.0)
