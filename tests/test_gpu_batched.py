"""Batched worlds: one device step over many independent worlds must leave
every world exactly as stepping it alone does (bitwise), and a world stepped
alone must match the oracle.  Needs an MI355X."""
import numpy as np
import pytest

from vxfix import OracleWorld, bits_equal, body_state

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ctx():
    from velox_amd.device import Context
    c = Context()
    yield c
    c.close()


def _worlds(name="batch", count=12):
    from velox_amd.device import Scene
    return Scene(name, count)


def test_batched_scene_layout():
    sc = _worlds()
    assert sc.worlds == 12
    counts = np.bincount(sc.world_of, minlength=sc.worlds)
    assert counts.min() >= 9 and counts.max() <= 513  # plane + 8..512 bodies
    assert (sc.bodies["shape"][sc.world_of == 3] == 1).sum() == 1  # one plane per world
    # Capsule chains: ball joints between consecutive capsules of one world.
    j = sc.joints
    assert len(j) > 0 and (j["type"] == 0).all()
    assert (sc.world_of[j["a"]] == sc.world_of[j["b"]]).all() and (j["b"] == j["a"] + 1).all()
    assert (sc.bodies["shape"][j["a"]] == 3).all() and (sc.bodies["shape"][j["b"]] == 3).all()
    sub = sc.world(int(sc.world_of[j["a"][0]]))
    assert len(sub.joints) > 0 and (sub.bodies["shape"][sub.joints["a"]] == 3).all()


@pytest.mark.parametrize("steps", [1, 25])
def test_batched_equals_each_world_alone(ctx, steps):
    from velox_amd.device import Context
    sc = _worlds()
    assert len(sc.joints), "the scene has capsule chains"
    print(f"batched 12 worlds, {len(sc.bodies)} bodies, {len(sc.joints)} joints, {steps} steps", flush=True)
    ctx.load_scene(sc)
    ctx.step(sc.dt, steps)
    batched = ctx.bodies()
    alone = Context()
    try:
        for w in range(sc.worlds):
            sub = sc.world(w)
            alone.load_scene(sub)
            alone.step(sc.dt, steps)
            got = batched[sc.world_of == w]
            assert bits_equal(body_state(got), body_state(alone.bodies())), f"world {w} differs after {steps} steps"
    finally:
        alone.close()


def test_batched_large_equals_each_world_alone(ctx):
    """256 worlds: the contact set is large enough for the device-wide colouring
    path, each world alone takes the single-workgroup one."""
    from velox_amd.device import Context
    sc = _worlds(count=256)
    steps = 3
    print(f"batched 256 worlds, {len(sc.bodies)} bodies, {steps} steps", flush=True)
    ctx.load_scene(sc)
    ctx.step(sc.dt, steps)
    assert ctx.stats()["contacts"] > (1 << 14)
    batched = ctx.bodies()
    alone = Context()
    try:
        for w in range(0, sc.worlds, 8):
            sub = sc.world(w)
            alone.load_scene(sub)
            alone.step(sc.dt, steps)
            got = batched[sc.world_of == w]
            assert bits_equal(body_state(got), body_state(alone.bodies())), f"world {w} differs after {steps} steps"
    finally:
        alone.close()


def test_batched_world_matches_oracle(ctx):
    """Worlds without capsule chains (the oracle has no joints; the chains'
    ball joints follow the reference World's device joint path, pinned by the
    joints fixtures)."""
    sc = _worlds()
    steps = 10
    ctx.load_scene(sc)
    ctx.step(sc.dt, steps)
    batched = ctx.bodies()
    jointed = set(sc.world_of[sc.joints["a"]].tolist())
    free = [w for w in range(sc.worlds) if w not in jointed][:2]
    assert len(free) == 2
    for w in free:
        sub = sc.world(w)
        ow = OracleWorld(sub.bodies, sub.soup, dt=sc.dt, substeps=sc.substeps, gravity=sc.gravity)
        ow.step(steps)
        got = batched[sc.world_of == w]
        assert bits_equal(body_state(got), body_state(ow.bodies)), f"world {w} differs from the oracle"


def test_batched_full_size_invariants(ctx):
    sc = _worlds(count=4096)
    print(f"batched 4096 worlds, {len(sc.bodies)} bodies", flush=True)
    ctx.load_scene(sc)
    ctx.step(sc.dt, 3)
    st = ctx.stats()
    b = ctx.bodies()
    assert np.isfinite(np.ascontiguousarray(body_state(b)).view(np.float32)).all()
    assert st["contacts"] > 0 and st["pairs"] > 0
    # No world moved a body of another world out of its own region: every
    # dynamic body is still above its world's plane.
    dyn = b["shape"] != 1
    assert (b["position"][dyn, 1] > -0.5).all()


def test_batched_restore_repeats_the_same_steps(ctx):
    """A saved state of the batched world (every world's bodies and carried
    contacts) replays the same steps after a restore: the batched snapshot
    the differentiable checkpoints need."""
    sc = _worlds()
    print(f"batched restore, {len(sc.bodies)} bodies", flush=True)
    ctx.load_scene(sc)
    ctx.step(sc.dt, 10)
    blob = ctx.save_state()
    ctx.step(sc.dt, 15)
    first = body_state(ctx.bodies())
    ctx.restore_state(blob)
    assert ctx.save_state() == blob
    ctx.step(sc.dt, 15)
    assert bits_equal(first, body_state(ctx.bodies()))
