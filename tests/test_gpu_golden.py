"""HIP device step against the reference golden vectors (bitwise), through the
C ABI of libvxhip.  Needs an MI355X."""
import numpy as np
import pytest

from vxfix import as_bodies, as_events, as_contacts, as_state, bits_equal, body_state, contact_bytes, default_options, load_case

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ctx():
    from velox_amd.device import Context
    c = Context()
    yield c
    c.close()


def _report(got, want):
    a8 = np.ascontiguousarray(got).view(np.uint8).reshape(len(got), -1)
    b8 = np.ascontiguousarray(want).view(np.uint8).reshape(len(want), -1)
    bad = np.nonzero((a8 != b8).any(axis=1))[0]
    return f"{len(bad)} of {len(got)} differ; first {bad[:8].tolist()}"


def test_integration_matches_reference(ctx):
    case = load_case("integrate")
    bodies = as_bodies(case["bodies_in"])
    ctx.integrate(bodies, (0.0, -9.81, 0.0), np.float32(1.0 / 240.0))
    want = as_bodies(case["after_integrate"])
    assert bits_equal(bodies, want), _report(bodies, want)


@pytest.mark.parametrize("name", ["freeflight", "lattice64", "lattice64_cone", "pile", "terrain", "lattice1000"])
def test_trajectory_matches_reference(ctx, name):
    case = load_case(name)
    meta = case["meta"]
    opts = default_options(friction_model=meta["frictionModel"], velocity_iterations=meta["velocityIterations"],
                           position_iterations=meta["positionIterations"])
    print(f"[gpu] trajectory {name}", flush=True)
    ctx.load(as_bodies(case["bodies_in"]), case["soup"], gravity=tuple(meta["gravity"]), substeps=meta["substeps"],
             options=opts)
    done = 0
    for k, target in enumerate(meta["steps"]):
        ctx.step(np.float32(meta["dt"]), target - done)
        done = target
        st = ctx.stats()
        got = body_state(ctx.bodies())
        want = as_state(case[f"state_{target}"])
        assert bits_equal(got, want), f"{name} step {target}: " + _report(got, want)
        assert st["contacts"] == meta["contacts_at"][k]
        assert st["colors"] == meta["colors"][k]
        if f"events_{target}" in case:
            want_e = as_events(case[f"events_{target}"])
            got_e = ctx.events()
            assert len(got_e) == len(want_e), f"{name} step {target}: {len(got_e)} events, want {len(want_e)}"
            assert np.array_equal(got_e.view(np.uint8), np.ascontiguousarray(want_e).view(np.uint8)), \
                f"{name} step {target} events"
        if f"contacts_{target}" in case:
            want_c = as_contacts(case[f"contacts_{target}"])
            assert np.array_equal(contact_bytes(ctx.contacts()), contact_bytes(want_c))


NARROW = ["nlattice_0", "nlattice_45", "npile_40", "nterrain_30", "nslab_2"]


@pytest.mark.parametrize("name", NARROW)
def test_broadphase_pairs_match_reference(ctx, name):
    case = load_case(name)
    bodies = as_bodies(case["bodies_in"])
    got = ctx.probe_pairs(bodies, case["soup"], np.float32(case["meta"]["dt"]))
    a, b = got >> np.uint64(32), got & np.uint64(0xffffffff)
    canon = (np.minimum(a, b) << np.uint64(32)) | np.maximum(a, b)
    assert np.array_equal(np.sort(canon), case["pairs"])
    assert np.all(np.diff(got.astype(np.float64)) > 0)  # sorted, unique contact keys


@pytest.mark.parametrize("name", NARROW)
def test_contacts_match_reference(ctx, name):
    case = load_case(name)
    bodies = as_bodies(case["bodies_in"])
    got = ctx.find_contacts(bodies, case["soup"], np.float32(case["meta"]["dt"]))
    want = as_contacts(case["contacts"])
    assert len(got) == len(want)
    assert np.array_equal(contact_bytes(got), contact_bytes(want))


@pytest.mark.parametrize("name", NARROW)
def test_coloring_matches_reference(ctx, name):
    case = load_case(name)
    colors, k = ctx.color(as_bodies(case["bodies_in"]), as_contacts(case["contacts"]))
    assert k == case["meta"]["colors"]
    assert np.array_equal(colors.astype(np.int32), case["colors"])


@pytest.mark.parametrize("name", NARROW)
def test_colored_solve_matches_reference(ctx, name):
    case = load_case(name)
    bodies = as_bodies(case["bodies_in"])
    cs = as_contacts(case["contacts"])
    h = np.float32(case["meta"]["dt"] / case["meta"]["substeps"])
    ctx.solve_velocities(bodies, cs, h, True)
    assert bits_equal(body_state(bodies), as_state(case["solve_bodies"]))
    ctx.fetch_impulses(cs)  # the solved impulses stay on the device until fetched
    imp = np.stack([cs["point"][:, 0], cs["point"][:, 1], cs["point"][:, 2], cs["normalImpulse"],
                    cs["tangentImpulse1"], cs["tangentImpulse2"], cs["rollingImpulse1"], cs["rollingImpulse2"],
                    cs["spinningImpulse"]], axis=1).astype(np.float32)
    assert bits_equal(imp, case["solve_impulses"])
