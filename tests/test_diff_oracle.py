"""The double-precision reference of the differentiable form (oracle/
vxo_diff.cpp: frozen-active-set rollout and its reverse, the CPU reference
adjoint) against central differences of its own forward, on a small world
whose frozen contacts are written out by hand: two cubes on a plane, one
pushed into the other, four corner contacts each with the plane and four
face contacts between them, the same contacts every step (each step's
contact k starts from the previous step's contact k).  This pins the
composition of the pieces tests/adjoint_check.cpp checks one by one: row
order, substeps, the warm start, the impulse chain across steps, the
positional correction and the loss."""
import ctypes as C

import numpy as np
import pytest

from vxfix import BODY, CONTACT, oracle, ptr

H = 0.25  # cube half-extent


def _cube(pos, vel, ang, mass, mu):
    b = np.zeros(1, BODY)[0]
    b["position"] = pos
    b["orientation"] = (0.0, 0.0, 0.0, 1.0)
    b["velocity"] = vel
    b["angularVelocity"] = ang
    b["invMass"] = 1.0 / mass
    b["invInertia"] = (6.0 / (mass * (2 * H) ** 2),) * 3
    b["inertiaOrientation"] = (0.0, 0.0, 0.0, 1.0)
    b["friction"] = mu
    b["frictionCombine"] = 1
    b["motionType"] = 2
    b["shape"] = 2
    b["halfExtents"] = (H, H, H)
    return b


def _world(params):
    """Plane + cube A at the origin + cube B touching A's +x face."""
    (vA, wA, vB, wB, muA, muB, mA, mB) = params
    bodies = np.zeros(3, BODY)
    bodies[0]["orientation"] = (0.0, 0.0, 0.0, 1.0)
    bodies[0]["friction"] = 0.6
    bodies[0]["frictionCombine"] = 1
    bodies[0]["motionType"] = 0
    bodies[0]["shape"] = 1
    bodies[0]["planeNormal"] = (0.0, 1.0, 0.0)
    gap = 5e-4
    bodies[1] = _cube((0.0, H + gap, 0.0), vA, wA, mA, muA)
    bodies[2] = _cube((2 * H + gap, H + gap, 0.0), vB, wB, mB, muB)
    cs = []
    for body in (1, 2):
        for sx in (-1, 1):
            for sz in (-1, 1):
                c = np.zeros(1, CONTACT)[0]
                local = np.array([sx * H, -H, sz * H], np.float32)
                world = bodies[body]["position"] + local
                c["a"], c["b"] = body, 0
                c["normal"] = (0.0, 1.0, 0.0)
                c["point"] = world
                c["localAnchorA"] = local
                c["localAnchorB"] = (world[0], 0.0, world[2])
                c["gap"] = c["bias0"] = world[1]
                mu = np.sqrt(np.float32(bodies[body]["friction"])) * np.sqrt(np.float32(0.6))
                c["friction1"] = c["friction2"] = mu
                cs.append(c)
    for sy in (-1, 1):
        for sz in (-1, 1):
            c = np.zeros(1, CONTACT)[0]
            local = np.array([H, sy * H, sz * H], np.float32)
            world = bodies[1]["position"] + local
            c["a"], c["b"] = 1, 2
            c["normal"] = (-1.0, 0.0, 0.0)
            c["point"] = world
            c["localAnchorA"] = local
            c["localAnchorB"] = world - bodies[2]["position"] + np.array([gap, 0, 0], np.float32)
            c["gap"] = c["bias0"] = gap
            mu = np.sqrt(np.float32(muA)) * np.sqrt(np.float32(muB))
            c["friction1"] = c["friction2"] = mu
            cs.append(c)
    return bodies, np.array(cs, CONTACT)


def _pack(params):
    """8 doubles per body (v0, w0, friction, mass); the plane's are unused."""
    (vA, wA, vB, wB, muA, muB, mA, mB) = params
    p = np.zeros((3, 8), np.float64)
    p[1, :3], p[1, 3:6], p[1, 6], p[1, 7] = vA, wA, muA, mA
    p[2, :3], p[2, 3:6], p[2, 6], p[2, 7] = vB, wB, muB, mB
    return p.reshape(-1)


def _rollout(params, steps, grad=False, base=None):
    bodies, cs = _world(base if base is not None else params)
    contacts = np.concatenate([cs] * steps)
    m = len(cs)
    match = np.concatenate([np.full(m, -1, np.int32)] + [np.arange(m, dtype=np.int32)] * (steps - 1))
    offsets = (np.arange(steps + 1) * m).astype(np.uint32)
    g = (C.c_float * 3)(0.0, -9.81, 0.0)
    target = (C.c_double * 3)(0.8, 0.3, -0.2)
    out = np.zeros(8 * len(bodies), np.float64) if grad else None
    lib = oracle()
    lib.vxo_diff_rollout.restype = C.c_double
    lib.vxo_diff_rollout.argtypes = [C.c_void_p, C.c_uint32, C.c_void_p, C.c_void_p, C.c_void_p, C.c_uint32, C.c_float,
                                     C.c_int, C.c_int, C.c_int, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p,
                                     C.c_void_p]
    packed = _pack(params)
    loss = lib.vxo_diff_rollout(ptr(bodies), len(bodies), ptr(contacts), ptr(match), ptr(offsets), steps,
                                np.float32(1.0 / 60.0), 4, 8, 3, g, target, ptr(packed), ptr(out) if grad else None,
                                None)
    return loss, out


@pytest.mark.parametrize("seed", [0, 1, 2])
def test_reference_adjoint_matches_central_differences(seed):
    rng = np.random.default_rng(seed)
    base = [np.array([rng.uniform(0.5, 2.0), 0.0, rng.uniform(-0.3, 0.3)]),
            np.array([0.0, rng.uniform(-0.5, 0.5), 0.0]),
            np.array([rng.uniform(-0.3, 0.0), 0.0, rng.uniform(-0.2, 0.2)]),
            np.array([0.0, rng.uniform(-0.3, 0.3), 0.0]),
            rng.uniform(0.2, 1.0), rng.uniform(0.2, 1.0), rng.uniform(0.5, 5.0), rng.uniform(0.5, 5.0)]
    steps = 12
    _, g = _rollout(base, steps, grad=True)
    g = g.reshape(3, 8)
    # (parameter slot, component) -> gradient entry of body 1 or 2
    checks = [(0, 0, 1, 0), (0, 2, 1, 2), (1, 1, 1, 4), (2, 0, 2, 0), (3, 1, 2, 4), (4, None, 1, 6),
              (5, None, 2, 6), (6, None, 1, 7), (7, None, 2, 7)]
    h = 1e-6
    got, want = [], []
    for slot, comp, body, col in checks:
        plus = [p.copy() if isinstance(p, np.ndarray) else p for p in base]
        minus = [p.copy() if isinstance(p, np.ndarray) else p for p in base]
        if comp is None:
            plus[slot] += h
            minus[slot] -= h
        else:
            plus[slot][comp] += h
            minus[slot][comp] -= h
        fd = (_rollout(plus, steps, base=base)[0] - _rollout(minus, steps, base=base)[0]) / (2 * h)
        got.append(g[body, col])
        want.append(fd)
    got, want = np.array(got), np.array(want)
    err = np.linalg.norm(got - want) / np.linalg.norm(want)
    assert err < 1e-4, (err, got, want)
