"""Shared test helpers: record dtypes, golden-fixture loading, the oracle's
ctypes binding and synthetic scenes.

Record layouts follow include/vx_records.h (byte-compatible with velox::Body,
velox::Contact, velox::Mesh, velox::BvhNode, velox::CompoundChild).  The
oracle (oracle/_build/liboracle.so) is the CPU checker; it is loaded here and
nowhere in the product path.
"""
import ctypes as C
import json
import os

import numpy as np

from velox_amd.records import (BODY, CONTACT, CONTACT_DATA_BYTES, BODY_DATA_BYTES, STATE, Soup,  # noqa: F401
                               SolverOptions, SoupArrays, default_options, ptr)

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
GOLDEN = os.path.join(ROOT, "tests", "golden")


def load_case(name):
    d = os.path.join(GOLDEN, name)
    case = {"dir": d, "meta": {}}
    if os.path.exists(os.path.join(d, "meta.json")):
        with open(os.path.join(d, "meta.json")) as f:
            case["meta"] = json.load(f)
    soup = {}
    for fn in sorted(os.listdir(d)):
        if not fn.endswith(".npy"):
            continue
        a = np.load(os.path.join(d, fn), allow_pickle=False)
        key = fn[:-4]
        if key.startswith("soup_"):
            soup[key[5:]] = a
        else:
            case[key] = a
    case["soup"] = SoupArrays(soup)
    return case


def as_bodies(raw):
    return np.ascontiguousarray(raw).view(BODY).reshape(-1).copy()


def as_contacts(raw):
    return np.ascontiguousarray(raw).view(CONTACT).reshape(-1).copy()


# vx_contact_event (include/vx_records.h), 44 B
EVENT = np.dtype([("a", np.uint32), ("b", np.uint32), ("point", np.float32, 3), ("normal", np.float32, 3),
                  ("impulse", np.float32), ("type", np.uint32), ("sensor", np.uint32)])
assert EVENT.itemsize == 44


def as_events(raw):
    return np.ascontiguousarray(raw).view(EVENT).reshape(-1)


def body_state(bodies):
    s = np.zeros(len(bodies), STATE)
    s["p"], s["q"], s["v"], s["w"] = bodies["position"], bodies["orientation"], bodies["velocity"], \
        bodies["angularVelocity"]
    return s


def as_state(raw):
    return np.ascontiguousarray(raw).view(STATE).reshape(-1)


def bits_equal(a, b):
    """Bitwise equality of two float/structured arrays (treats -0 != +0, NaN payloads exact)."""
    return np.array_equal(np.ascontiguousarray(a).view(np.uint8), np.ascontiguousarray(b).view(np.uint8))


def contact_bytes(cs):
    return np.ascontiguousarray(cs).view(np.uint8).reshape(-1, 128)[:, :CONTACT_DATA_BYTES]


# --------------------------------------------------------------- oracle --
class OracleConfig(C.Structure):
    _fields_ = [("gravity", C.c_float * 3), ("substeps", C.c_int32), ("options", SolverOptions)]


class OracleStats(C.Structure):
    _fields_ = [("pairs", C.c_uint32), ("contacts", C.c_uint32), ("colors", C.c_uint32),
                ("ccdRescued", C.c_uint32)]


_oracle = None


def oracle():
    """Load the CPU oracle (test infrastructure)."""
    global _oracle
    if _oracle is None:
        path = os.path.join(ROOT, "oracle", "_build", "liboracle.so")
        lib = C.CDLL(path)
        vp, u32, f32, i32 = C.c_void_p, C.c_uint32, C.c_float, C.c_int
        sig = {
            "vxo_integrate": (None, [vp, u32, vp, f32]),
            "vxo_advance_transforms": (None, [vp, u32, f32]),
            "vxo_body_aabbs": (None, [vp, u32, vp, f32, vp]),
            "vxo_broadphase": (u32, [vp, u32, vp, f32, vp, u32]),
            "vxo_collide_pair": (i32, [vp, vp, u32, u32, f32, vp]),
            "vxo_sphere_pair_cpu": (i32, [vp, u32, u32, f32, vp]),
            "vxo_find_contacts": (u32, [vp, u32, vp, f32, vp, u32]),
            "vxo_warm_start_match": (None, [vp, u32, vp, u32]),
            "vxo_std_sort": (None, [vp, vp, u32]),
            "vxo_diff_rollout": (C.c_double, [vp, u32, vp, vp, vp, u32, f32, i32, i32, i32, vp, vp, vp, vp, vp]),
            "vxo_color": (u32, [vp, u32, vp, u32, vp]),
            "vxo_solve": (None, [vp, vp, u32, f32, i32, i32, vp]),
            "vxo_position_correction": (None, [vp, vp, u32, f32, vp]),
            "vxo_ccd": (u32, [vp, vp, vp, vp, u32, vp, f32]),
            "vxo_gap_at": (f32, [vp, u32, u32, vp, f32, vp]),
            "vxo_world_step": (i32, [vp, u32, vp, vp, f32, vp, vp, u32, vp, u32, vp]),
            "vxo_contact_events": (u32, [vp, u32, vp, u32, f32, vp, vp, u32, vp, u32, vp, u32]),
        }
        for name, (res, args) in sig.items():
            fn = getattr(lib, name)
            fn.restype, fn.argtypes = res, args
        _oracle = lib
    return _oracle


class OracleWorld:
    """Sequential CPU replay of the device step on one world."""

    def __init__(self, bodies, soup, dt=1.0 / 60.0, substeps=4, gravity=(0.0, -9.81, 0.0), options=None,
                 contact_cap=None):
        self.bodies = np.ascontiguousarray(bodies).copy()
        self.soup = soup
        self._view = soup.view()
        self.dt = dt
        self.cfg = OracleConfig((C.c_float * 3)(*gravity), substeps, options or default_options())
        cap = contact_cap or max(4096, 64 * len(self.bodies))
        self.prev = np.zeros(cap, CONTACT)
        self.prev_count = C.c_uint32(0)
        self.contacts = np.zeros(cap, CONTACT)
        self.stats = OracleStats()
        self.last_contacts = self.contacts[:0]
        self.event_keys = np.zeros(cap, np.uint64)
        self.event_key_count = C.c_uint32(0)
        self.events = np.zeros(2 * cap, EVENT)
        self.last_events = self.events[:0]

    def step(self, n=1):
        lib = oracle()
        for _ in range(n):
            prev_contacts = self.prev[:self.prev_count.value].copy()
            m = lib.vxo_world_step(ptr(self.bodies), len(self.bodies), C.byref(self._view), C.byref(self.cfg),
                                   self.dt, ptr(self.prev), C.byref(self.prev_count), len(self.prev),
                                   ptr(self.contacts), len(self.contacts), C.byref(self.stats))
            if m > len(self.contacts):
                raise RuntimeError("oracle contact capacity exceeded")
            self.last_contacts = self.contacts[:m].copy()
            ne = lib.vxo_contact_events(ptr(self.bodies), len(self.bodies), ptr(self.last_contacts), m, self.dt,
                                        ptr(self.event_keys), C.byref(self.event_key_count), len(self.event_keys),
                                        ptr(prev_contacts), len(prev_contacts), ptr(self.events), len(self.events))
            if ne > len(self.events):
                raise RuntimeError("oracle event capacity exceeded")
            self.last_events = self.events[:ne].copy()
        return self
