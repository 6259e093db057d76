"""Sphere-sphere contacts: the reference CPU backend's analytic contact
(collideSpherePair, solver.cpp:26-52, taken by CpuBackend::findContacts,
solver.cpp:478-483) next to the generic collidePair that the device backend
runs for every pair (narrowphase.h; backend_cuda.cu).

Fixture tests/golden/spheres, written by oracle/ref_harness.cpp from the
strict reference library: 712 spheres (uniform cluster, eight pairs with
coincident centres, 64 pairs placed within a few float steps of the analytic
path's reach), the exact-AABB pair set in canonical orientation, the CPU
backend's contacts for those pairs (contacts_cpu) and the generic ones
(contacts).

What the fixture shows, and these tests pin: wherever the analytic path emits
a contact it is byte-identical to the generic one (the reference's claim,
solver.cpp:20-23); but the analytic path returns no contact for a pair beyond
(|va| + |vb|) dt + 1e-3 (with a 1e-5 relative band), while the generic emit()
keeps speculative contacts up to that reach plus both bodies' speculative
distances (narrowphase.h:425-441).  The device follows the device backend
(generic path), so it reports those extra speculative sphere contacts too.
"""
import numpy as np
import pytest

from vxfix import as_bodies, as_contacts, contact_bytes, load_case, oracle, ptr, CONTACT


def _pair_contacts(name="spheres"):
    case = load_case(name)
    bodies = as_bodies(case["bodies_in"])
    return case, bodies, as_contacts(case["contacts"]), as_contacts(case["contacts_cpu"])


def _keys(cs):
    return ((cs["a"].astype(np.uint64) << np.uint64(32)) | cs["b"].astype(np.uint64)).tolist()


def test_oracle_restates_cpu_backend_sphere_contacts():
    """The oracle's restatement of collideSpherePair, with the generic path
    for the pairs it defers (coincident centres), reproduces the CPU backend's
    contacts byte for byte, in pair order."""
    case, bodies, generic, cpu = _pair_contacts()
    lib = oracle()
    out = np.zeros(16, CONTACT)
    got, deferred, beyond = [], 0, 0
    dt = np.float32(case["meta"]["dt"])
    for key in case["pairs"].tolist():
        i, j = key >> 32, key & 0xffffffff
        assert bodies["shape"][i] == 0 and bodies["shape"][j] == 0  # spheres only
        n = lib.vxo_sphere_pair_cpu(ptr(bodies), i, j, dt, ptr(out))
        if n < 0:
            deferred += 1
            n = lib.vxo_collide_pair(ptr(bodies), None, i, j, dt, ptr(out))
        elif n == 0:
            beyond += 1
        got.extend(out[:n].copy())
    got = np.array(got, CONTACT)
    assert deferred == case["meta"]["coincident"] == 8
    assert beyond > 0
    assert len(got) == len(cpu)
    assert np.array_equal(contact_bytes(got), contact_bytes(cpu))


def test_analytic_contacts_equal_generic_ones_where_both_emit():
    case, bodies, generic, cpu = _pair_contacts()
    kg, kc = _keys(generic), _keys(cpu)
    assert len(set(kg)) == len(kg) and len(set(kc)) == len(kc)  # one contact per sphere pair
    at = {k: i for i, k in enumerate(kg)}
    assert set(kc) <= set(kg)
    gb, cb = contact_bytes(generic), contact_bytes(cpu)
    for i, k in enumerate(kc):
        assert np.array_equal(cb[i], gb[at[k]]), f"pair {k >> 32},{k & 0xffffffff}"


def test_generic_only_contacts_lie_between_the_two_reaches():
    """Each contact only the generic path emits belongs to a pair the analytic
    path leaves out for distance, and lies within emit()'s reach."""
    case, bodies, generic, cpu = _pair_contacts()
    lib = oracle()
    out = np.zeros(16, CONTACT)
    dt = np.float32(case["meta"]["dt"])
    extra = sorted(set(_keys(generic)) - set(_keys(cpu)))
    assert len(extra) > 100
    spec = bodies["speculativeDistance"]
    at = {k: i for i, k in enumerate(_keys(generic))}
    for k in extra:
        i, j = k >> 32, k & 0xffffffff
        assert lib.vxo_sphere_pair_cpu(ptr(bodies), i, j, dt, ptr(out)) == 0
        c = generic[at[k]]
        speed = np.linalg.norm(bodies["velocity"][[i, j]].astype(np.float64), axis=1).sum()
        dist = np.linalg.norm(bodies["position"][i].astype(np.float64) - bodies["position"][j])
        reach_analytic = bodies["radius"][i] + bodies["radius"][j] + speed * float(dt) + 1e-3
        assert dist > reach_analytic * (1 - 1e-4)
        extra_reach = 1e-3 + speed * float(dt) + float(spec[i]) + float(spec[j])
        assert c["gap"] <= extra_reach * (1 + 1e-4)


@pytest.mark.gpu
def test_device_contacts_match_generic_and_cpu_backend():
    """The device reports the generic path's contacts byte for byte and, on
    every pair the CPU backend emits one, that backend's contact exactly."""
    from velox_amd.device import Context
    case, bodies, generic, cpu = _pair_contacts()
    ctx = Context()
    try:
        got = ctx.find_contacts(bodies, case["soup"], np.float32(case["meta"]["dt"]))
    finally:
        ctx.close()
    assert len(got) == len(generic)
    assert np.array_equal(contact_bytes(got), contact_bytes(generic))
    at = {k: i for i, k in enumerate(_keys(got))}
    gb, cb = contact_bytes(got), contact_bytes(cpu)
    for i, k in enumerate(_keys(cpu)):
        assert np.array_equal(cb[i], gb[at[k]])
