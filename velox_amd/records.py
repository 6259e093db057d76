"""Record layouts of the device step (include/vx_records.h) as numpy dtypes and
ctypes structures.  Byte-compatible with velox::Body (320 B), velox::Contact
(128 B), velox::Mesh, velox::BvhNode, velox::CompoundChild and
velox::SolverOptions (include/velox/body.h, backend.h, solver.h)."""
import ctypes as C

import numpy as np


def _dtype(fields, itemsize):
    """Structured dtype whose gaps are filled with void fields: numpy does not
    copy the bytes of unnamed holes, and record bytes are compared bitwise."""
    full, at = [], 0
    for name, fmt, off in sorted(fields, key=lambda f: f[2]):
        if off > at:
            full.append((f"_pad{at}", f"V{off - at}", at))
        full.append((name, fmt, off))
        at = off + np.dtype(fmt).itemsize
    if at < itemsize:
        full.append((f"_pad{at}", f"V{itemsize - at}", at))
    names, formats, offsets = zip(*full)
    return np.dtype({"names": list(names), "formats": list(formats), "offsets": list(offsets),
                     "itemsize": itemsize})


BODY = _dtype([
    ("position", ("<f4", 3), 0), ("orientation", ("<f4", 4), 12), ("velocity", ("<f4", 3), 28),
    ("angularVelocity", ("<f4", 3), 40), ("force", ("<f4", 3), 52), ("torque", ("<f4", 3), 64),
    ("invMass", "<f4", 76), ("invInertia", ("<f4", 3), 80), ("inertiaOrientation", ("<f4", 4), 92),
    ("restitution", "<f4", 108), ("friction", "<f4", 112), ("frictionScale", ("<f4", 3), 116),
    ("rollingFriction", "<f4", 128), ("spinningFriction", "<f4", 132), ("frictionCombine", "u1", 136),
    ("restitutionCombine", "u1", 137), ("linearDamping", "<f4", 140), ("angularDamping", "<f4", 144),
    ("gravityScale", "<f4", 148), ("motionType", "u1", 152), ("ccdQuality", "u1", 156),
    ("collisionMargin", "<f4", 160), ("speculativeDistance", "<f4", 164), ("enableContinuous", "u1", 168),
    ("minVelocityForCCD", "<f4", 172), ("geometryQuality", "u1", 208), ("categoryBits", "<u4", 212),
    ("maskBits", "<u4", 216), ("groupIndex", "<i4", 220), ("sensor", "u1", 224), ("enableSleep", "u1", 225),
    ("fixedRotation", "u1", 226), ("asleep", "u1", 227), ("sleepTimer", "<f4", 228), ("shape", "u1", 232),
    ("radius", "<f4", 236), ("halfExtents", ("<f4", 3), 240), ("capsuleHalfHeight", "<f4", 252),
    ("planeNormal", ("<f4", 3), 256), ("planeOffset", "<f4", 268), ("meshIndex", "<u4", 272),
    ("hullFirst", "<u4", 276), ("hullCount", "<u4", 280), ("hullFaceFirst", "<u4", 284),
    ("hullFaceCount", "<u4", 288), ("compoundFirst", "<u4", 292), ("compoundCount", "<u4", 296),
], 320)

CONTACT = _dtype([
    ("a", "<u4", 0), ("b", "<u4", 4), ("featureKey", "<u8", 8), ("normal", ("<f4", 3), 16),
    ("point", ("<f4", 3), 28), ("localAnchorA", ("<f4", 3), 40), ("localAnchorB", ("<f4", 3), 52),
    ("gap", "<f4", 64), ("bias0", "<f4", 68), ("vn0", "<f4", 72), ("restitution", "<f4", 76),
    ("friction1", "<f4", 80), ("friction2", "<f4", 84), ("rollingFriction", "<f4", 88),
    ("spinningFriction", "<f4", 92), ("normalImpulse", "<f4", 96), ("tangentImpulse1", "<f4", 100),
    ("tangentImpulse2", "<f4", 104), ("rollingImpulse1", "<f4", 108), ("rollingImpulse2", "<f4", 112),
    ("spinningImpulse", "<f4", 116),
], 128)

# Contact bytes that carry data (the record's tail is alignment padding).
CONTACT_DATA_BYTES = 120
BODY_DATA_BYTES = 300

STATE = np.dtype([("p", "<f4", 3), ("q", "<f4", 4), ("v", "<f4", 3), ("w", "<f4", 3)])

SHAPE_SPHERE, SHAPE_PLANE, SHAPE_BOX, SHAPE_CAPSULE, SHAPE_MESH, SHAPE_HULL = 0, 1, 2, 3, 4, 5


class Soup(C.Structure):
    """vx_soup: raw pointers plus element counts (include/vx_records.h)."""
    _fields_ = [
        ("vertices", C.c_void_p), ("vertexCount", C.c_uint32),
        ("indices", C.c_void_p), ("indexCount", C.c_uint32),
        ("meshes", C.c_void_p), ("meshCount", C.c_uint32),
        ("bvhNodes", C.c_void_p), ("bvhNodeCount", C.c_uint32),
        ("bvhTriRefs", C.c_void_p), ("bvhTriRefCount", C.c_uint32),
        ("hullPoints", C.c_void_p), ("hullPointCount", C.c_uint32),
        ("hullFaceIndices", C.c_void_p), ("hullFaceIndexCount", C.c_uint32),
        ("compoundChildren", C.c_void_p), ("compoundChildCount", C.c_uint32),
    ]


class SolverOptions(C.Structure):
    _fields_ = [("impulseThreshold", C.c_float), ("stackStiffness", C.c_float), ("stackDamping", C.c_float),
                ("velocityIterations", C.c_int32), ("positionIterations", C.c_int32),
                ("frictionModel", C.c_uint8), ("iterationPolicy", C.c_uint8),
                ("enableStackStabilization", C.c_uint8), ("reserved", C.c_uint8)]


def default_options(friction_model=0, velocity_iterations=8, position_iterations=3):
    return SolverOptions(1e-5, 1000.0, 10.0, velocity_iterations, position_iterations, friction_model, 0, 0, 0)



def ptr(a):
    """Raw pointer of a contiguous numpy array (None when empty)."""
    return None if a is None or a.size == 0 else C.c_void_p(a.ctypes.data)


# vx_contact_event (include/vx_records.h), 44 B
EVENT = np.dtype([("a", np.uint32), ("b", np.uint32), ("point", np.float32, 3), ("normal", np.float32, 3),
                  ("impulse", np.float32), ("type", np.uint32), ("sensor", np.uint32)])
assert EVENT.itemsize == 44


class SoupArrays:
    """Owns the numpy arrays behind one vx_soup view."""

    SLOTS = [("vertices", np.float32, 3), ("indices", np.uint32, 0), ("meshes", np.uint8, 52),
             ("nodes", np.uint8, 32), ("trirefs", np.uint32, 0), ("hullpts", np.float32, 3),
             ("hullfaces", np.uint32, 0), ("children", np.uint8, 68)]

    def __init__(self, arrays=None):
        arrays = arrays or {}
        for name, dt, width in self.SLOTS:
            a = arrays.get(name)
            if a is None:
                a = np.zeros((0, width) if width else (0,), dt)
            setattr(self, name, np.ascontiguousarray(a, dtype=dt))

    def view(self):
        return Soup(ptr(self.vertices), len(self.vertices), ptr(self.indices), len(self.indices),
                    ptr(self.meshes), len(self.meshes), ptr(self.nodes), len(self.nodes),
                    ptr(self.trirefs), len(self.trirefs), ptr(self.hullpts), len(self.hullpts),
                    ptr(self.hullfaces), len(self.hullfaces), ptr(self.children), len(self.children))


# vx_joint (include/vx_records.h): velox::Joint, 420 B.  Only the fields the
# host reads by name are spelled out; the rest are named raw-byte fields, so
# that numpy copies (which skip unnamed bytes) keep the whole record.
JOINT = np.dtype([("type", "u1"), ("_pad1", "V3"), ("a", "<u4"), ("b", "<u4"), ("_config", "V380"),
                  ("reactionLinearImpulse", "<f4", 3), ("reactionAngularImpulse", "<f4", 3), ("broken", "u1"),
                  ("_pad417", "V3")])
assert JOINT.itemsize == 420 and JOINT.fields["reactionLinearImpulse"][1] == 392 and JOINT.fields["broken"][1] == 416

# vx_joint_break (include/vxhip.h), 20 B.
JOINT_BREAK = np.dtype([("index", "<u4"), ("a", "<u4"), ("b", "<u4"), ("force", "<f4"), ("torque", "<f4")])

# vx_toi_hit (include/vxhip.h), 36 B: MultiToiHit with the dense body index.
TOI_HIT = np.dtype([("toi", "<f4"), ("body", "<u4"), ("point", "<f4", 3), ("normal", "<f4", 3), ("fraction", "<f4")])
assert TOI_HIT.itemsize == 36
