"""Device step against the oracle at workload scale (bitwise), and
size-independent invariants at the full 50,000-body size.  Needs an MI355X."""
import numpy as np
import pytest

from vxfix import CONTACT, OracleWorld, bits_equal, body_state

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


@pytest.mark.parametrize("name,param,steps,every", [("lattice", 10, 30, 10), ("pile", 3000, 40, 10),
                                                    ("debris", 50000, 2, 1)])
def test_device_step_matches_oracle(ctx, name, param, steps, every):
    from velox_amd.device import Scene
    sc = Scene(name, param)
    print(f"[gpu] {name} {param}: {len(sc.bodies)} bodies, {steps} steps", flush=True)
    ctx.load_scene(sc)
    ref = OracleWorld(sc.bodies, sc.soup, dt=np.float32(sc.dt), substeps=sc.substeps, gravity=sc.gravity,
                      contact_cap=max(1 << 16, 24 * len(sc.bodies)))
    for s in range(1, steps + 1):
        ctx.step(np.float32(sc.dt))
        ref.step()
        if s % every == 0 or s == steps:
            st = ctx.stats()
            assert st["contacts"] == ref.stats.contacts, (name, s)
            assert st["colors"] == ref.stats.colors, (name, s)
            assert st["pairs"] == ref.stats.pairs, (name, s)
            got = body_state(ctx.bodies())
            want = body_state(ref.bodies)
            assert bits_equal(got, want), f"{name} step {s}: " + _report(got, want)
            got_e, want_e = ctx.events(), ref.last_events
            assert len(got_e) == len(want_e), (name, s, len(got_e), len(want_e))
            assert np.array_equal(got_e.view(np.uint8), np.ascontiguousarray(want_e).view(np.uint8)), (name, s)
    print(f"[gpu] {name}: {ctx.stats()}", flush=True)


@pytest.mark.parametrize("limit", [4, 64])
def test_big_island_path_matches_oracle(ctx, limit):
    """Islands above `limit` contacts are solved colour by colour over the
    whole device; the trajectory must stay bitwise equal to the oracle."""
    from velox_amd.device import Context, Scene
    sc = Scene("pile", 600)
    steps = 30
    print(f"pile 600, island limit {limit}, {steps} steps", flush=True)
    c = Context()
    try:
        c.load_scene(sc)
        c.set_island_limit(limit)
        ow = OracleWorld(sc.bodies, sc.soup, dt=sc.dt, substeps=sc.substeps, gravity=sc.gravity)
        for step in range(steps):
            c.step(sc.dt, 1)
            ow.step(1)
            got = c.bodies()
            assert bits_equal(body_state(got), body_state(ow.bodies)), \
                f"step {step + 1}: {_report(body_state(got), body_state(ow.bodies))}"
        assert c.stats()["bigIslandContacts"] > 0
    finally:
        c.close()


def test_full_size_invariants(ctx):
    """50,000 bodies, 60 steps: finite state, no body below the plane by more
    than the speculative slop, colouring valid (no two contacts of a colour
    share a dynamic body), contacts sorted by pair key within the canonical
    order the solver replays."""
    from velox_amd.device import Scene
    sc = Scene("debris", 50000)
    ctx.load_scene(sc)
    ctx.step(np.float32(sc.dt), 60)
    b = ctx.bodies()
    assert np.isfinite(b["position"]).all() and np.isfinite(b["velocity"]).all()
    dyn = b["motionType"] == 2
    lowest = b["position"][dyn, 1] - np.where(b["shape"][dyn] == 0, b["radius"][dyn], 0)
    assert lowest[b["shape"][dyn] == 0].min() > -0.05
    cs = ctx.contacts()
    st = ctx.stats()
    assert len(cs) == st["contacts"] > 0
    # colour classes are contiguous in solve order; rebuild them and check independence
    dynamic = dyn
    seen = {}
    from velox_amd.device import Context  # noqa: F401
    colors, k = ctx.color(b, np.sort(cs, order=["a", "b"]))
    assert k >= 1
    order = np.sort(cs, order=["a", "b"])
    for col in range(k):
        sel = order[colors == col]
        ends = np.concatenate([sel["a"][dynamic[sel["a"]]], sel["b"][dynamic[sel["b"]]]])
        assert len(np.unique(ends)) == len(ends), f"colour {col} shares a dynamic body"
    del seen
