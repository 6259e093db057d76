"""Pins the CPU oracle against golden vectors written by the reference Velox
library itself (strict floating point; tests/golden/make_golden.py).

Every comparison here is bitwise: the oracle restates the reference's
arithmetic in the same operation order, so any dropped term, swapped operand
or reordered sum shows up as a bit difference.
"""
import ctypes as C

import numpy as np
import pytest

from vxfix import (as_events, CONTACT, OracleWorld, as_bodies, as_contacts, as_state, bits_equal, body_state,
                   contact_bytes, default_options, load_case, oracle, ptr)


def _diff_report(a, b):
    a8 = np.ascontiguousarray(a).view(np.uint8).reshape(len(a), -1)
    b8 = np.ascontiguousarray(b).view(np.uint8).reshape(len(b), -1)
    bad = np.nonzero((a8 != b8).any(axis=1))[0]
    return f"{len(bad)} of {len(a)} records differ; first {bad[:8].tolist()}"


def test_integration_matches_reference():
    case = load_case("integrate")
    bodies = as_bodies(case["bodies_in"])
    lib = oracle()
    g = (C.c_float * 3)(0.0, -9.81, 0.0)
    h = np.float32(1.0 / 240.0)
    lib.vxo_integrate(ptr(bodies), len(bodies), g, h)
    want = as_bodies(case["after_integrate"])
    assert bits_equal(bodies, want), _diff_report(bodies, want)
    lib.vxo_advance_transforms(ptr(bodies), len(bodies), h)
    want = as_bodies(case["after_transform"])
    assert bits_equal(bodies, want), _diff_report(bodies, want)
    for _ in range(99):
        lib.vxo_integrate(ptr(bodies), len(bodies), g, h)
        lib.vxo_advance_transforms(ptr(bodies), len(bodies), h)
    want = as_bodies(case["after_100"])
    assert bits_equal(bodies, want), _diff_report(bodies, want)


def _run_trajectory(name, cone=False):
    case = load_case(name)
    meta = case["meta"]
    opts = default_options(friction_model=meta["frictionModel"], velocity_iterations=meta["velocityIterations"],
                           position_iterations=meta["positionIterations"])
    w = OracleWorld(as_bodies(case["bodies_in"]), case["soup"], dt=np.float32(meta["dt"]),
                    substeps=meta["substeps"], gravity=tuple(meta["gravity"]), options=opts)
    done = 0
    for k, target in enumerate(meta["steps"]):
        w.step(target - done)
        done = target
        got = body_state(w.bodies)
        want = as_state(case[f"state_{target}"])
        assert bits_equal(got, want), f"{name} step {target}: " + _diff_report(got, want)
        assert w.stats.contacts == meta["contacts_at"][k], (name, target)
        assert w.stats.colors == meta["colors"][k], (name, target)
        if f"events_{target}" in case:
            want_e = as_events(case[f"events_{target}"])
            assert np.array_equal(np.ascontiguousarray(w.last_events).view(np.uint8),
                                  np.ascontiguousarray(want_e).view(np.uint8)), f"{name} step {target} events"
        if f"contacts_{target}" in case:
            want_c = as_contacts(case[f"contacts_{target}"])
            got_c = w.last_contacts
            assert len(got_c) == len(want_c)
            assert np.array_equal(contact_bytes(got_c), contact_bytes(want_c)), f"{name} step {target} contacts"
    if "bodies_out" in case:
        assert bits_equal(w.bodies, as_bodies(case["bodies_out"]))


def test_free_flight_trajectory_matches_reference():
    _run_trajectory("freeflight")


def test_box_lattice_trajectory_matches_reference():
    _run_trajectory("lattice64")


def test_crowded_slab_stack_matches_reference():
    """One 35,000-contact island and 700+ colours: the colour spill past one
    64-bit mask word and the big-island regime of a large settled pile."""
    _run_trajectory("slabstack")


def test_box_lattice_cone_block_trajectory_matches_reference():
    _run_trajectory("lattice64_cone")


def test_mixed_pile_trajectory_matches_reference():
    _run_trajectory("pile")


def test_terrain_trajectory_matches_reference():
    _run_trajectory("terrain")


NARROW = ["nlattice_0", "nlattice_45", "npile_40", "nterrain_30", "nslab_2"]


@pytest.mark.parametrize("name", NARROW)
def test_broadphase_aabbs_and_pairs_match_reference(name):
    case = load_case(name)
    bodies = as_bodies(case["bodies_in"])
    dt = np.float32(case["meta"]["dt"])
    lib = oracle()
    view = case["soup"].view()
    lohi = np.zeros((len(bodies), 6), np.float32)
    lib.vxo_body_aabbs(ptr(bodies), len(bodies), C.byref(view), dt, ptr(lohi))
    assert bits_equal(lohi, case["aabbs"])
    pairs = np.zeros(len(case["pairs"]) + 16, np.uint64)
    n = lib.vxo_broadphase(ptr(bodies), len(bodies), C.byref(view), dt, ptr(pairs), len(pairs))
    # The oracle lists pairs by contact key (plane/mesh pairs name the convex
    # body first); the reference lists (lo<<32|hi).  Same set either way.
    got = pairs[:n]
    lo = np.minimum(got >> np.uint64(32), got & np.uint64(0xffffffff))
    hi = np.maximum(got >> np.uint64(32), got & np.uint64(0xffffffff))
    assert np.array_equal(np.sort((lo << np.uint64(32)) | hi), case["pairs"])


@pytest.mark.parametrize("name", NARROW)
def test_narrowphase_contacts_match_reference(name):
    case = load_case(name)
    bodies = as_bodies(case["bodies_in"])
    dt = np.float32(case["meta"]["dt"])
    want = as_contacts(case["contacts"])
    out = np.zeros(len(want) + 64, CONTACT)
    n = oracle().vxo_find_contacts(ptr(bodies), len(bodies), C.byref(case["soup"].view()), dt, ptr(out), len(out))
    assert n == len(want)
    assert np.array_equal(contact_bytes(out[:n]), contact_bytes(want))


@pytest.mark.parametrize("name", NARROW)
def test_coloring_and_colored_sweeps_match_reference(name):
    case = load_case(name)
    bodies = as_bodies(case["bodies_in"])
    cs = as_contacts(case["contacts"])
    lib = oracle()
    colors = np.zeros(len(cs), np.uint32)
    K = lib.vxo_color(ptr(bodies), len(bodies), ptr(cs), len(cs), ptr(colors))
    assert K == case["meta"]["colors"]
    assert np.array_equal(colors.astype(np.int32), case["colors"])
    order = np.argsort(colors, kind="stable")
    solved = cs[order].copy()
    h = np.float32(case["meta"]["dt"] / case["meta"]["substeps"])
    opts = default_options()
    lib.vxo_solve(ptr(bodies), ptr(solved), len(solved), h, 1, 16, C.byref(opts))
    assert bits_equal(body_state(bodies), as_state(case["solve_bodies"]))
    imp = np.stack([solved["point"][:, 0], solved["point"][:, 1], solved["point"][:, 2], solved["normalImpulse"],
                    solved["tangentImpulse1"], solved["tangentImpulse2"], solved["rollingImpulse1"],
                    solved["rollingImpulse2"], solved["spinningImpulse"]], axis=1).astype(np.float32)
    assert bits_equal(imp, case["solve_impulses"])
