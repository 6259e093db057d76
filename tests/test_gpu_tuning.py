"""Scheduling knobs (vxhip_tuning): every setting must step a world bit for
bit the same.  Covers the persistent longest-first island-batch sweeps, the
512-thread batch workgroups, the colour-launch threshold, the island limit,
the joint-pack limit (packs of one workgroup against a launch per colour), the
side stream and the hipGraph-captured substeps (each graph replay checked
against eagerly launched substeps, step by step, with the program's own syncs
and downloads between the replays).  Needs an MI355X."""
import numpy as np
import pytest

from vxfix import OracleWorld, as_bodies, bits_equal, body_state, load_case

pytestmark = pytest.mark.gpu

DEFAULTS = dict(batchThreads=256, persistentBatches=True, longestFirst=True, colorLaunchRows=1024,
                islandLimit=4096, sideStream=True, graphSubsteps=False, jointPackMax=4096)


@pytest.fixture(scope="module")
def ctx():
    from velox_amd.device import Context
    c = Context()
    yield c
    c.close()


def _trajectory(ctx, scene, knobs, steps):
    ctx.load_scene(scene, tuned=False)
    ctx.set_tuning(**knobs)
    states = []
    for _ in range(steps):
        ctx.step(scene.dt, 1)
        states.append(body_state(ctx.bodies()))
    return states


def test_tuning_round_trip(ctx):
    ctx.set_tuning(batchThreads=512, graphSubsteps=True, sideStream=False, colorLaunchRows=700)
    t = ctx.tuning()
    assert t["batchThreads"] == 512 and t["graphSubsteps"] and not t["sideStream"] and t["colorLaunchRows"] == 700
    assert t["persistentBatches"] and t["longestFirst"] and t["islandLimit"] == 4096
    ctx.set_tuning()
    assert ctx.tuning() == DEFAULTS
    from velox_amd.device import VxhipError
    with pytest.raises(VxhipError):
        ctx.set_tuning(batchThreads=384)


@pytest.mark.parametrize("knobs", [
    dict(batchThreads=512),
    dict(persistentBatches=False, longestFirst=False),
    dict(longestFirst=False),
    dict(graphSubsteps=True),
    dict(graphSubsteps=True, batchThreads=512, persistentBatches=False),
    dict(jointPackMax=1),
])
def test_batched_worlds_any_knobs_bitwise(ctx, knobs):
    """256 batched worlds, 12 steps: each knob setting against the defaults,
    every step."""
    from velox_amd.device import Scene
    sc = Scene("batch", 256)
    print(f"batched 256 worlds, knobs {knobs}", flush=True)
    want = _trajectory(ctx, sc, DEFAULTS, 12)
    got = _trajectory(ctx, sc, dict(DEFAULTS, **knobs), 12)
    for t, (g, w) in enumerate(zip(got, want)):
        assert bits_equal(g, w), f"step {t + 1}: knobs {knobs} change the state"
    st = ctx.stats()
    if knobs.get("graphSubsteps"):
        assert st["graphCaptures"] > 0 and st["graphReplays"] > 0, st


def test_graph_substeps_match_oracle(ctx):
    """The lattice64 fixture with graph-captured substeps: every step bitwise
    against the oracle's replay, and the graph really replayed."""
    case = load_case("lattice64")
    bodies = as_bodies(case["bodies_in"])
    ctx.set_tuning(**dict(DEFAULTS, graphSubsteps=True))
    ctx.load(bodies, case["soup"])
    ref = OracleWorld(bodies, case["soup"], dt=np.float32(1.0 / 60.0))
    before = ctx.stats()["graphReplays"]
    for t in range(20):
        ctx.step(np.float32(1.0 / 60.0))
        ref.step()
        ctx.sync()
        assert bits_equal(ctx.bodies(), ref.bodies), f"step {t + 1} differs from the oracle"
    st = ctx.stats()
    assert st["graphReplays"] - before >= 20, st
    ctx.set_tuning()


@pytest.mark.parametrize("knobs", [
    dict(colorLaunchRows=256),
    dict(colorLaunchRows=4096, islandLimit=2048),
    dict(islandLimit=16384),
    dict(sideStream=False),
    dict(batchThreads=512, persistentBatches=False),
])
def test_big_island_any_knobs_bitwise(ctx, knobs):
    """The crowded slab stack (one island of ~35k contacts, hundreds of
    colours, island batches beside it): knob settings against the defaults."""
    from velox_amd.records import default_options
    case = load_case("slabstack")
    meta = case["meta"]
    bodies = as_bodies(case["bodies_in"])
    print(f"slabstack, knobs {knobs}", flush=True)
    out = []
    for k in (DEFAULTS, dict(DEFAULTS, **knobs)):
        ctx.set_tuning(**k)
        ctx.load(bodies, case["soup"], gravity=tuple(meta["gravity"]), substeps=meta["substeps"],
                 options=default_options())
        ctx.step(np.float32(meta["dt"]), 2)
        out.append(body_state(ctx.bodies()))
        st = ctx.stats()
        assert st["bigIslandContacts"] > 0
    ctx.set_tuning()
    assert bits_equal(out[1], out[0]), f"knobs {knobs} change the slab stack"
