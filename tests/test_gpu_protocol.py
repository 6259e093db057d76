"""The device backend's residency protocol through the C ABI (Backend::
solveVelocities / fetchImpulses / invalidateCaches, backend_cuda.cu:569-690,
763-783), against the oracle's coloured replay.  Needs an MI355X.

  * a warm-started solve returns the caller's contacts sorted and in colour
    order with the impulses they came in with; the solved impulses stay on the
    device until fetched;
  * a later solve with warmStart = 0 reuses the device-resident contacts and
    replays the sweeps with the step and options of the warm-started call
    (the reference replays the graph it captured there), whatever it is given;
  * fetchImpulses returns the impulses after every solve since;
  * invalidateCaches drops the device-resident contact set (a cold solve is
    refused, fetch leaves the caller's array alone) and the resident world's
    warm-start cache (its next step starts cold, as the oracle world does with
    an empty previous-contact list);
  * a Backend-equivalent call detaches a resident world (stepping it again is
    refused until it is loaded again).
"""
import ctypes as C

import numpy as np
import pytest

from vxfix import (OracleWorld, as_bodies, as_contacts, as_state, bits_equal, body_state, contact_bytes,
                   default_options, load_case, oracle, ptr)

pytestmark = pytest.mark.gpu

IMPULSES = ("normalImpulse", "tangentImpulse1", "tangentImpulse2", "rollingImpulse1", "rollingImpulse2",
            "spinningImpulse")


@pytest.fixture(scope="module")
def ctx():
    from velox_amd.device import Context
    c = Context()
    yield c
    c.close()


def _impulses(cs):
    return np.stack([cs[f] for f in IMPULSES], 1).astype(np.float32)


def _integrate_on_host(bodies, h):
    """The caller's work between substeps (integration and transforms), done
    by the oracle on both copies alike."""
    lib = oracle()
    g = (C.c_float * 3)(0.0, -9.81, 0.0)
    lib.vxo_integrate(ptr(bodies), len(bodies), g, h)
    lib.vxo_advance_transforms(ptr(bodies), len(bodies), h)


@pytest.mark.parametrize("name", ["npile_40", "nlattice_45"])
def test_warm_then_cold_solve_and_fetch_match_oracle(ctx, name):
    case = load_case(name)
    h = np.float32(case["meta"]["dt"] / case["meta"]["substeps"])
    opts = default_options()
    sweeps = 2 * opts.velocityIterations
    bodies_d = as_bodies(case["bodies_in"])
    cs_d = as_contacts(case["contacts"])
    # the oracle: canonical contacts in colour order (the fixture's colours)
    bodies_o = as_bodies(case["bodies_in"])
    order = np.argsort(case["colors"], kind="stable")
    cs_o = as_contacts(case["contacts"])[order].copy()
    print(f"[gpu] protocol {name}: {len(cs_d)} contacts", flush=True)

    ctx.solve_velocities(bodies_d, cs_d, h, True, opts)
    assert ctx.last_velocity_iterations() == sweeps
    assert np.array_equal(contact_bytes(cs_d), contact_bytes(cs_o))  # colour order, input impulses
    oracle().vxo_solve(ptr(bodies_o), ptr(cs_o), len(cs_o), h, 1, sweeps, C.byref(opts))
    assert bits_equal(body_state(bodies_d), body_state(bodies_o))

    _integrate_on_host(bodies_d, h)
    _integrate_on_host(bodies_o, h)
    # a different step and iteration count on the cold call: the warm call's are replayed
    other = default_options(velocity_iterations=3)
    ctx.solve_velocities(bodies_d, cs_d, np.float32(h * 0.5), False, other)
    oracle().vxo_solve(ptr(bodies_o), ptr(cs_o), len(cs_o), h, 0, sweeps, C.byref(opts))
    assert bits_equal(body_state(bodies_d), body_state(bodies_o))
    assert ctx.last_velocity_iterations() == 0  # reported by the warm-started call only

    before = _impulses(cs_d).copy()
    ctx.fetch_impulses(cs_d)
    assert bits_equal(_impulses(cs_d), _impulses(cs_o))
    assert not bits_equal(_impulses(cs_d), before)
    assert np.array_equal(cs_d["a"], cs_o["a"]) and np.array_equal(cs_d["b"], cs_o["b"])


def test_invalidate_drops_the_device_contact_set(ctx):
    from velox_amd.device import VxhipError
    case = load_case("npile_40")
    h = np.float32(case["meta"]["dt"] / case["meta"]["substeps"])
    bodies = as_bodies(case["bodies_in"])
    cs = as_contacts(case["contacts"])
    ctx.solve_velocities(bodies, cs, h, True)
    ctx.invalidate_caches()
    mark = cs.copy()
    ctx.fetch_impulses(cs)  # nothing solved since the invalidation: left alone
    assert np.array_equal(contact_bytes(cs), contact_bytes(mark))
    with pytest.raises(VxhipError):
        ctx.solve_velocities(bodies, cs, h, False)
    ctx.solve_velocities(bodies, cs, h, True)  # a warm-started solve works again


def test_fetch_with_another_count_leaves_contacts_alone(ctx):
    case = load_case("nlattice_45")
    h = np.float32(case["meta"]["dt"] / case["meta"]["substeps"])
    bodies = as_bodies(case["bodies_in"])
    cs = as_contacts(case["contacts"])
    ctx.solve_velocities(bodies, cs, h, True)
    short = cs[:-1].copy()
    ctx.fetch_impulses(short)
    assert np.array_equal(contact_bytes(short), contact_bytes(cs[:-1]))


@pytest.mark.parametrize("name", ["lattice64", "pile"])
def test_invalidate_makes_the_resident_world_start_cold(ctx, name):
    case = load_case(name)
    meta = case["meta"]
    opts = default_options(friction_model=meta["frictionModel"], velocity_iterations=meta["velocityIterations"],
                           position_iterations=meta["positionIterations"])
    bodies = as_bodies(case["bodies_in"])
    dt = np.float32(meta["dt"])
    ctx.load(bodies, case["soup"], gravity=tuple(meta["gravity"]), substeps=meta["substeps"], options=opts)
    ref = OracleWorld(bodies, case["soup"], dt=dt, substeps=meta["substeps"], gravity=tuple(meta["gravity"]),
                      options=opts)
    ctx.step(dt, 5)
    ref.step(5)
    ctx.invalidate_caches()
    ref.prev_count.value = 0  # an empty previous-contact list: the next step starts cold
    ctx.step(dt, 2)
    ref.step(2)
    assert bits_equal(body_state(ctx.bodies()), body_state(ref.bodies))
    # and it differs from the warm-started continuation (the test is not vacuous)
    ctx.load(bodies, case["soup"], gravity=tuple(meta["gravity"]), substeps=meta["substeps"], options=opts)
    ctx.step(dt, 7)
    assert not bits_equal(body_state(ctx.bodies()), body_state(ref.bodies))


def test_backend_call_detaches_the_resident_world(ctx):
    from velox_amd.device import VxhipError
    case = load_case("lattice64")
    meta = case["meta"]
    bodies = as_bodies(case["bodies_in"])
    ctx.load(bodies, case["soup"], gravity=tuple(meta["gravity"]), substeps=meta["substeps"])
    ctx.step(np.float32(meta["dt"]), 1)
    ctx.integrate(bodies.copy(), (0.0, -9.81, 0.0), np.float32(1.0 / 240.0))
    with pytest.raises(VxhipError):
        ctx.step(np.float32(meta["dt"]), 1)
    ctx.load(bodies, case["soup"], gravity=tuple(meta["gravity"]), substeps=meta["substeps"])
    ctx.step(np.float32(meta["dt"]), 1)
