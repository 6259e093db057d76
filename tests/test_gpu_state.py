"""Origin shift and state save/restore of the resident device world.

`shift_lattice64` and `shift_terrain` are written by the reference World
(oracle/ref_harness.cpp, `fixture shift`): a 4x4x4 box lattice shifted by
(1000.25, -2.5, 500.75) after 30 steps and stepped to 60, and a heightfield
terrain scene (mesh vertices and BVH bounds move too) shifted by
(-64.5, 3.125, 17) after 20 steps and stepped to 40.  Bodies, the last events
and the warm-start cache right after the shift, and states and events at the
end, are compared bit for bit.  Save/restore is checked for determinism:
stepping after a restore repeats the steps after the save exactly.
Needs an MI355X."""
import numpy as np
import pytest

from vxfix import as_bodies, as_contacts, as_events, as_state, bits_equal, body_state, default_options, \
    load_case

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ctx():
    from velox_amd.device import Context
    c = Context()
    yield c
    c.close()


def _load(ctx, case):
    meta = case["meta"]
    opts = default_options(friction_model=meta["frictionModel"], velocity_iterations=meta["velocityIterations"],
                           position_iterations=meta["positionIterations"])
    ctx.load(as_bodies(case["bodies_in"]), case["soup"], gravity=tuple(meta["gravity"]), substeps=meta["substeps"],
             options=opts)
    return meta


def _cache_from_blob(blob, n):
    """Warm-start cache records out of a saved state (header, bodies, cache)."""
    hdr = np.frombuffer(blob[:32], np.uint32)
    assert hdr[2] == n
    start = 32 + n * 320
    return np.frombuffer(blob[start:start + int(hdr[3]) * 128], np.uint8).reshape(-1, 128).copy()


def _sorted_rows(rows):
    rows = np.ascontiguousarray(rows)
    return rows[np.lexsort(rows.T[::-1])]


@pytest.mark.parametrize("name", ["shift_lattice64", "shift_terrain"])
def test_origin_shift_matches_reference(ctx, name):
    case = load_case(name)
    meta = _load(ctx, case)
    dt = np.float32(meta["dt"])
    print(f"[gpu] {name} shift at {meta['shift_at']}", flush=True)
    ctx.step(dt, meta["shift_at"])
    ctx.shift_origin(meta["offset"])
    got_b = ctx.bodies()
    want_b = as_bodies(case["bodies_shifted"])
    assert bits_equal(body_state(got_b), body_state(want_b)), "bodies after the shift"
    assert bits_equal(got_b["planeOffset"], want_b["planeOffset"]), "plane offsets after the shift"
    assert bits_equal(ctx.events(), as_events(case["events_shifted"])), "events after the shift"
    # The cache holds the same records as the reference's (order within a
    # pair is the reference's std::sort order there, canonical here).
    cache = _cache_from_blob(ctx.save_state(), len(got_b))
    want_c = np.ascontiguousarray(as_contacts(case["prev_shifted"])).view(np.uint8).reshape(-1, 128)
    from vxfix import CONTACT_DATA_BYTES
    assert np.array_equal(_sorted_rows(cache[:, :CONTACT_DATA_BYTES]), _sorted_rows(want_c[:, :CONTACT_DATA_BYTES]))
    end = meta["steps"][-1]
    ctx.step(dt, end - meta["shift_at"])
    assert bits_equal(body_state(ctx.bodies()), as_state(case[f"state_{end}"])), f"states at step {end}"
    assert bits_equal(ctx.events(), as_events(case[f"events_{end}"])), f"events at step {end}"


def test_origin_shift_overflow_changes_nothing(ctx):
    case = load_case("shift_lattice64")
    meta = _load(ctx, case)
    ctx.step(np.float32(meta["dt"]), 5)
    ctx.shift_origin((-3.0e38, 0.0, 0.0))
    before = ctx.save_state()
    from velox_amd.device import VxhipError
    with pytest.raises(VxhipError):
        ctx.shift_origin((-3.0e38, 0.0, 0.0))  # positions would pass FLT_MAX
    with pytest.raises(VxhipError):
        ctx.shift_origin((float("nan"), 0.0, 0.0))
    assert ctx.save_state() == before


@pytest.mark.parametrize("name,at,more", [("lattice64", 10, 25), ("pile", 5, 30), ("shift_terrain", 12, 20)])
def test_restore_repeats_the_same_steps(ctx, name, at, more):
    case = load_case(name)
    meta = _load(ctx, case)
    dt = np.float32(meta["dt"])
    print(f"[gpu] restore {name} at {at}", flush=True)
    ctx.step(dt, at)
    blob = ctx.save_state()
    ctx.step(dt, more)
    first = (body_state(ctx.bodies()), ctx.events(), ctx.stats()["contacts"])
    ctx.restore_state(blob)
    assert ctx.save_state() == blob
    ctx.step(dt, more)
    second = (body_state(ctx.bodies()), ctx.events(), ctx.stats()["contacts"])
    assert bits_equal(first[0], second[0]) and bits_equal(first[1], second[1]) and first[2] == second[2]


def test_restore_rejects_foreign_state(ctx):
    from velox_amd.device import VxhipError
    case = load_case("lattice64")
    _load(ctx, case)
    blob = bytearray(ctx.save_state())
    with pytest.raises(VxhipError):
        ctx.restore_state(bytes(blob[:-1]))
    blob[0] ^= 0xFF
    with pytest.raises(VxhipError):
        ctx.restore_state(bytes(blob))
