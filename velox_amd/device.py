"""ctypes binding of libvxhip (include/vxhip.h), the MI355X device step.

The library is built in-tree (velox_amd/_build/libvxhip.so) by
velox_amd.build.build(); loading fails loudly when it is missing or when no
gfx950 device is present -- there is no CPU fallback.
"""
import ctypes as C
import os

import numpy as np

from .records import BODY, CONTACT, SolverOptions, default_options, ptr

HERE = os.path.dirname(os.path.abspath(__file__))
LIBRARY = os.path.join(HERE, "_build", "libvxhip.so")

OK, ERR_MEMORY, ERR_DEVICE_LOST, ERR_INVALID, ERR_CAPACITY, ERR_UNAVAILABLE = range(6)


class VxhipError(RuntimeError):
    def __init__(self, code, message):
        super().__init__(f"vxhip error {code}: {message}")
        self.code = code


class Config(C.Structure):
    _fields_ = [("gravity", C.c_float * 3), ("substeps", C.c_int32), ("options", SolverOptions)]


class Stats(C.Structure):
    _fields_ = [("bodies", C.c_uint32), ("pairs", C.c_uint32), ("contacts", C.c_uint32), ("colors", C.c_uint32),
                ("ccdCandidates", C.c_uint32), ("ccdRescued", C.c_uint32), ("velocityIterations", C.c_uint32),
                ("oversize", C.c_uint32), ("cellSize", C.c_float), ("islands", C.c_uint32),
                ("islandContacts", C.c_uint32), ("bigIslandContacts", C.c_uint32),
                ("largestIsland", C.c_uint32), ("maxBatchPasses", C.c_uint32), ("multiToiEvents", C.c_uint32),
                ("ccdFlagged", C.c_uint32), ("ccdRelevant", C.c_uint32), ("warmTies", C.c_uint32),
                ("launches", C.c_uint32 * 6), ("graphCaptures", C.c_uint32), ("graphReplays", C.c_uint32)]

    def as_dict(self):
        d = {name: getattr(self, name) for name, _ in self._fields_}
        d["launches"] = dict(zip(LAUNCH_KINDS, self.launches))
        return d


# vxhip_stats.launches indices (VXHIP_LAUNCH_*)
LAUNCH_KINDS = ("batch", "island", "color_pair", "color_run", "tail", "greedy_device")


class SleepConfig(C.Structure):
    """vx_sleep_config (SleepConfig, sleep.h:98-160) with the reference defaults."""
    _fields_ = [("linearVelocityThreshold", C.c_float), ("angularVelocityThreshold", C.c_float),
                ("accelerationThreshold", C.c_float), ("timeToSleep", C.c_float), ("timeToDrowsy", C.c_float),
                ("drowsySimulationRate", C.c_float), ("contactStabilityThreshold", C.c_float),
                ("contactStabilityFrames", C.c_int32), ("enableGradualSleep", C.c_uint8),
                ("enableContactStability", C.c_uint8), ("reserved", C.c_uint8 * 2)]

    def __init__(self, **kw):
        defaults = dict(linearVelocityThreshold=0.05, angularVelocityThreshold=0.05, accelerationThreshold=0.01,
                        timeToSleep=0.5, timeToDrowsy=0.25, drowsySimulationRate=0.25,
                        contactStabilityThreshold=0.02, contactStabilityFrames=10, enableGradualSleep=1,
                        enableContactStability=1)
        defaults.update(kw)
        super().__init__(**defaults)


class MultiToiSettings(C.Structure):
    """vx_multi_toi_settings (WorldMultiToiSettings, ccd.h:111-122) with the reference defaults."""
    _fields_ = [("enabled", C.c_uint8), ("enableSubstepSplitting", C.c_uint8), ("reserved", C.c_uint8 * 2),
                ("maxToiEventsPerBody", C.c_uint32), ("toiVelocityFloor", C.c_float),
                ("toiPenetrationBias", C.c_float), ("maxTotalEventsPerStep", C.c_uint32)]

    def __init__(self, **kw):
        defaults = dict(enabled=1, enableSubstepSplitting=1, maxToiEventsPerBody=4, toiVelocityFloor=0.1,
                        toiPenetrationBias=1e-3, maxTotalEventsPerStep=256)
        defaults.update(kw)
        super().__init__(**defaults)


class Tuning(C.Structure):
    """vxhip_tuning (include/vxhip.h): scheduling knobs, none of which changes
    a result.  Switch fields: 0 default, 1 on, 2 off."""
    _fields_ = [("islandLimit", C.c_uint32), ("batchThreads", C.c_uint32), ("persistentBatches", C.c_uint32),
                ("longestFirst", C.c_uint32), ("colorLaunchRows", C.c_uint32), ("sideStream", C.c_uint32),
                ("graphSubsteps", C.c_uint32), ("jointPackMax", C.c_uint32)]
    SWITCHES = ("persistentBatches", "longestFirst", "sideStream", "graphSubsteps")

    def as_dict(self):
        d = {name: getattr(self, name) for name, _ in self._fields_}
        for k in self.SWITCHES:
            d[k] = d[k] == 1
        return d

    @classmethod
    def from_dict(cls, d):
        t = cls()
        for name, _ in cls._fields_:
            if name in d:
                v = d[name]
                setattr(t, name, (1 if v else 2) if name in cls.SWITCHES else int(v))
        return t


class DiffStats(C.Structure):
    """vxhip_diff_stats (include/vxhip.h)."""
    _fields_ = [(name, C.c_uint32) for name in ("steps", "checkpoints", "worlds", "ccdRescued", "replayMismatch",
                                                "maxContacts")]

    def as_dict(self):
        return {name: getattr(self, name) for name, _ in self._fields_}


class SleepStats(C.Structure):
    _fields_ = [(name, C.c_uint32) for name in ("totalDynamicBodies", "awakeBodies", "drowsyBodies", "sleepingBodies",
                                                "islandCount", "sleepingIslandCount", "drowsyIslandCount",
                                                "awakeIslandCount")] + \
               [(name, C.c_uint64) for name in ("totalSleepTransitions", "totalWakeTransitions",
                                                "totalDrowsyTransitions")]

    def as_dict(self):
        return {name: getattr(self, name) for name, _ in self._fields_}


# Exported symbols and their signatures (include/vxhip.h).
_vp, _u32, _f32, _i32 = C.c_void_p, C.c_uint32, C.c_float, C.c_int
SIGNATURES = {
    "vxhip_device_count": (_i32, [_vp]),
    "vxhip_last_error": (C.c_char_p, []),
    "vxhip_create": (_i32, [_vp]),
    "vxhip_destroy": (None, [_vp]),
    "vxhip_integrate": (_i32, [_vp, _vp, _u32, _vp, _f32]),
    "vxhip_find_contacts": (_i32, [_vp, _vp, _u32, _vp, _f32, _vp, _u32, _vp]),
    "vxhip_solve_velocities": (_i32, [_vp, _vp, _u32, _vp, _u32, _f32, _i32, _vp]),
    "vxhip_advance_substeps": (_i32, [_vp, _vp, _u32, _vp, _u32, _vp, _u32, _vp, _f32, _i32, _vp, _vp]),
    "vxhip_fetch_impulses": (_i32, [_vp, _vp, _u32]),
    "vxhip_invalidate_caches": (_i32, [_vp]),
    "vxhip_last_velocity_iterations": (_u32, [_vp]),
    "vxhip_world_load": (_i32, [_vp, _vp, _u32, _vp, _vp]),
    "vxhip_world_step": (_i32, [_vp, _f32, _u32]),
    "vxhip_world_set_worlds": (_i32, [_vp, _vp, _u32]),
    "vxhip_world_set_island_limit": (_i32, [_vp, _u32]),
    "vxhip_world_set_tuning": (_i32, [_vp, _vp]),
    "vxhip_world_get_tuning": (_i32, [_vp, _vp]),
    "vxhip_world_download_events": (_i32, [_vp, _vp, _u32, _vp]),
    "vxhip_world_set_events": (_i32, [_vp, _i32]),
    "vxhip_world_set_sleep": (_i32, [_vp, _vp]),
    "vxhip_world_shift_origin": (_i32, [_vp, _vp]),
    "vxhip_world_set_multi_toi": (_i32, [_vp, _vp]),
    "vxhip_world_query_multi_toi": (_i32, [_vp, _u32, _f32, _vp, _u32, _vp]),
    "vxhip_world_set_joints": (_i32, [_vp, _vp, _u32]),
    "vxhip_world_set_joints_all": (_i32, [_vp, _vp, _u32]),
    "vxhip_world_download_joints": (_i32, [_vp, _vp, _u32, _vp]),
    "vxhip_world_download_joint_breaks": (_i32, [_vp, _vp, _u32, _vp]),
    "vxhip_world_state_size": (_i32, [_vp, _vp]),
    "vxhip_world_save_state": (_i32, [_vp, _vp, C.c_size_t]),
    "vxhip_world_restore_state": (_i32, [_vp, _vp, C.c_size_t]),
    "vxhip_world_sleep_stats": (_i32, [_vp, _vp]),
    "vxhip_world_sync": (_i32, [_vp]),
    "vxhip_world_download_bodies": (_i32, [_vp, _vp, _u32]),
    "vxhip_world_upload_bodies": (_i32, [_vp, _vp, _u32]),
    "vxhip_world_download_contacts": (_i32, [_vp, _vp, _u32, _vp]),
    "vxhip_world_stats": (_i32, [_vp, _vp]),
    "vxhip_probe_pairs": (_i32, [_vp, _vp, _u32, _vp, _f32, _vp, _u32, _vp]),
    "vxhip_probe_color": (_i32, [_vp, _vp, _u32, _vp, _u32, _vp, _vp]),
    "vxhip_probe_warm_start": (_i32, [_vp, _vp, _u32, _vp, _u32]),
    "vxhip_probe_std_sort": (_i32, [_vp, _vp, _vp, _u32, _vp, _u32]),
    "vxhip_probe_radix_sort": (_i32, [_vp, _vp, _vp, _u32, _i32]),
    "vxhip_std_sort_noalloc": (_i32, [_vp, _vp, _vp, _u32, _vp, _u32]),
    "vxhip_world_record": (_i32, [_vp, _i32]),
    "vxhip_world_download_record": (_i32, [_vp, _vp, _vp, _u32, _vp]),
    "vxhip_diff_forward": (_i32, [_vp, _f32, _u32, _u32, _vp, _vp]),
    "vxhip_diff_backward": (_i32, [_vp, _vp, _vp, _vp, _vp, _vp, _vp]),
    "vxhip_world_profile": (_i32, [_vp, _i32]),
    "vxhip_world_timings": (_i32, [_vp, _vp]),
    "vxhip_scene_build": (_i32, [C.c_char_p, _u32, _vp]),
    "vxhip_scene_body_count": (_u32, [_vp]),
    "vxhip_scene_bodies": (_vp, [_vp]),
    "vxhip_scene_worlds": (_vp, [_vp, _vp]),
    "vxhip_scene_soup": (_i32, [_vp, _vp]),
    "vxhip_scene_joints": (_vp, [_vp, _vp]),
    "vxhip_scene_config": (_i32, [_vp, _vp, _vp]),
    "vxhip_scene_free": (None, [_vp]),
}

_lib = None


def library():
    """Load libvxhip.so (raises if it has not been built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIBRARY):
            raise RuntimeError(f"{LIBRARY} is missing: run velox_amd.build.build() (or __graft_entry__.build())")
        # One HIP runtime per process: when PyTorch is installed it is loaded
        # first, and the library binds to the runtime PyTorch brings, so device
        # buffers and streams can pass between the two (the differentiable
        # form's autograd binding, caller-owned device buffers).
        try:
            import torch  # noqa: F401
        except ImportError:
            pass
        lib = C.CDLL(LIBRARY)
        for name, (res, args) in SIGNATURES.items():
            fn = getattr(lib, name)
            fn.restype, fn.argtypes = res, args
        _lib = lib
    return _lib


def check(code):
    if code != OK:
        raise VxhipError(code, library().vxhip_last_error().decode())


def device_count():
    n = C.c_int(0)
    check(library().vxhip_device_count(C.byref(n)))
    return n.value


class Timings(C.Structure):
    _fields_ = [("steps", C.c_uint32), ("beginMs", C.c_float), ("broadMs", C.c_float), ("narrowMs", C.c_float),
                ("matchMs", C.c_float), ("colorMs", C.c_float), ("solveMs", C.c_float), ("ccdMs", C.c_float),
                ("positionMs", C.c_float), ("finishMs", C.c_float), ("kernelMs", C.c_float * 8),
                ("kernelLaunches", C.c_uint32 * 8), ("kernelVisits", C.c_uint64 * 8)]

    # Event-timed kernels, in the order of the per-kernel arrays (VXHIP_KERNEL_*).
    KERNELS = ("kBatchSweeps<0>", "kBatchSweeps<2>", "kIslandSweeps", "kSweepColorPair/Run/TailPair", "kNarrow",
               "kCcdCheck", "kCcdDataflow", "kGridPairs")

    def as_dict(self):
        d = {}
        for name, _ in self._fields_:
            v = getattr(self, name)
            d[name] = list(v) if name.startswith("kernel") else v
        return d


class Scene:
    """Body records + flat geometry of a synthetic workload, built host-side by
    the library (vxhip_scene_build)."""

    def __init__(self, name, param=0):
        from .records import SoupArrays
        lib = library()
        h = C.c_void_p()
        self.name, self.param = name, int(param)
        check(lib.vxhip_scene_build(name.encode(), int(param), C.byref(h)))
        try:
            n = lib.vxhip_scene_body_count(h)
            raw = (C.c_uint8 * (320 * n)).from_address(lib.vxhip_scene_bodies(h)) if n else b""
            self.bodies = np.frombuffer(bytes(raw), np.uint8).reshape(-1, 320).copy().view(BODY).reshape(-1)
            wc = C.c_uint32(0)
            wp = lib.vxhip_scene_worlds(h, C.byref(wc))
            self.worlds = wc.value
            self.world_of = np.frombuffer(bytes((C.c_uint8 * (4 * n)).from_address(wp)), np.uint32).copy() \
                if n else np.zeros(0, np.uint32)
            from .records import JOINT
            jc = C.c_uint32(0)
            jp = lib.vxhip_scene_joints(h, C.byref(jc))
            self.joints = np.frombuffer(bytes((C.c_uint8 * (JOINT.itemsize * jc.value)).from_address(jp)),
                                        JOINT).copy() if jc.value else np.zeros(0, JOINT)
            view = SoupView()
            check(lib.vxhip_scene_soup(h, C.byref(view)))
            cfg = Config()
            dt = C.c_float(0)
            check(lib.vxhip_scene_config(h, C.byref(cfg), C.byref(dt)))
            self.gravity = tuple(cfg.gravity)
            self.substeps = cfg.substeps
            self.dt = dt.value

            def grab(p, count, dtype, width):
                if not count:
                    return None
                item = np.dtype(dtype).itemsize * (width or 1)
                a = np.frombuffer(bytes((C.c_uint8 * (item * count)).from_address(p)), dtype)
                return a.reshape(count, width).copy() if width else a.copy()
            self.soup = SoupArrays({
                "vertices": grab(view.vertices, view.vertexCount, np.float32, 3),
                "indices": grab(view.indices, view.indexCount, np.uint32, 0),
                "meshes": grab(view.meshes, view.meshCount, np.uint8, 52),
                "nodes": grab(view.bvhNodes, view.bvhNodeCount, np.uint8, 32),
                "trirefs": grab(view.bvhTriRefs, view.bvhTriRefCount, np.uint32, 0),
                "hullpts": grab(view.hullPoints, view.hullPointCount, np.float32, 3),
                "hullfaces": grab(view.hullFaceIndices, view.hullFaceIndexCount, np.uint32, 0),
            })
        finally:
            lib.vxhip_scene_free(h)

    def world(self, w):
        """World w of a batched scene as a scene of its own (same geometry
        arrays, the world's bodies in their batched order)."""
        sub = object.__new__(Scene)
        sub.name, sub.param = self.name, self.param
        mine = self.world_of == w
        sub.bodies = np.ascontiguousarray(self.bodies[mine])
        local = np.cumsum(mine, dtype=np.int64) - 1  # body index -> index within world w
        j = self.joints[mine[self.joints["a"]]] if len(self.joints) else self.joints
        sub.joints = j.copy()
        if len(j):
            sub.joints["a"] = local[j["a"]]
            sub.joints["b"] = local[j["b"]]
        sub.worlds = 1
        sub.world_of = np.zeros(len(sub.bodies), np.uint32)
        sub.gravity, sub.substeps, sub.dt, sub.soup = self.gravity, self.substeps, self.dt, self.soup
        return sub


class SoupView(C.Structure):
    _fields_ = [("vertices", C.c_void_p), ("vertexCount", C.c_uint32), ("indices", C.c_void_p),
                ("indexCount", C.c_uint32), ("meshes", C.c_void_p), ("meshCount", C.c_uint32),
                ("bvhNodes", C.c_void_p), ("bvhNodeCount", C.c_uint32), ("bvhTriRefs", C.c_void_p),
                ("bvhTriRefCount", C.c_uint32), ("hullPoints", C.c_void_p), ("hullPointCount", C.c_uint32),
                ("hullFaceIndices", C.c_void_p), ("hullFaceIndexCount", C.c_uint32),
                ("compoundChildren", C.c_void_p), ("compoundChildCount", C.c_uint32)]


class Context:
    """One vxhip context: a device-resident world plus the Backend-style stage
    entry points."""

    def __init__(self):
        lib = library()
        h = C.c_void_p()
        check(lib.vxhip_create(C.byref(h)))
        self._h = h
        self.n = 0

    def close(self):
        if self._h:
            library().vxhip_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # -- resident world --------------------------------------------------
    def load_scene(self, scene, options=None, tuned=True):
        """Loads a synthetic scene; with `tuned`, the knobs the autotuner's
        table (velox_amd/tuning.json) holds for the scene's workload are set
        first (results are the same with any knobs)."""
        if tuned:
            from .tuning import knobs_for
            self.set_tuning(**knobs_for(scene))
        self.load(scene.bodies, scene.soup, gravity=scene.gravity, substeps=scene.substeps, options=options)
        if getattr(scene, "worlds", 1) > 1:
            self.set_worlds(scene.world_of, scene.worlds)
        joints = getattr(scene, "joints", None)
        if joints is not None and len(joints):
            self.set_joints(joints, any_count=True)
        return self

    def set_tuning(self, **knobs):
        """Sets vxhip_tuning fields by name (unnamed fields take their defaults)."""
        t = Tuning.from_dict(knobs)
        check(library().vxhip_world_set_tuning(self._h, C.byref(t)))
        return self

    def tuning(self):
        t = Tuning()
        check(library().vxhip_world_get_tuning(self._h, C.byref(t)))
        return t.as_dict()

    def set_island_limit(self, limit):
        """Islands above `limit` contacts are solved colour by colour over the device."""
        check(library().vxhip_world_set_island_limit(self._h, int(limit)))
        return self

    def set_worlds(self, world_of, worlds):
        """Batched worlds: body i belongs to world world_of[i]."""
        world_of = np.ascontiguousarray(world_of, dtype=np.uint32)
        assert len(world_of) == self.n
        check(library().vxhip_world_set_worlds(self._h, ptr(world_of), int(worlds)))
        return self

    def load(self, bodies, soup, gravity=(0.0, -9.81, 0.0), substeps=4, options=None):
        self._soup = soup
        self._view = soup.view()
        cfg = Config((C.c_float * 3)(*gravity), substeps, options or default_options())
        bodies = np.ascontiguousarray(bodies)
        assert bodies.dtype == BODY
        check(library().vxhip_world_load(self._h, ptr(bodies), len(bodies), C.byref(self._view), C.byref(cfg)))
        self.n = len(bodies)
        return self

    def step(self, dt, steps=1):
        check(library().vxhip_world_step(self._h, float(dt), int(steps)))

    def sync(self):
        check(library().vxhip_world_sync(self._h))

    def bodies(self):
        out = np.zeros(self.n, BODY)
        check(library().vxhip_world_download_bodies(self._h, ptr(out), self.n))
        return out

    def upload_bodies(self, bodies):
        bodies = np.ascontiguousarray(bodies)
        check(library().vxhip_world_upload_bodies(self._h, ptr(bodies), len(bodies)))

    def contacts(self):
        st = self.stats()
        out = np.zeros(max(1, st["contacts"]), CONTACT)
        cnt = C.c_uint32(0)
        check(library().vxhip_world_download_contacts(self._h, ptr(out), len(out), C.byref(cnt)))
        return out[:cnt.value]

    def profile(self, enable=True):
        check(library().vxhip_world_profile(self._h, int(enable)))

    def timings(self):
        t = Timings()
        check(library().vxhip_world_timings(self._h, C.byref(t)))
        return t.as_dict()

    def events(self):
        """Contact events of the last step (vx_contact_event records)."""
        from .records import EVENT
        count = C.c_uint32(0)
        rc = library().vxhip_world_download_events(self._h, None, 0, C.byref(count))
        if rc not in (OK, ERR_CAPACITY):
            check(rc)
        out = np.zeros(count.value, EVENT)
        check(library().vxhip_world_download_events(self._h, ptr(out) if len(out) else None, len(out), C.byref(count)))
        return out

    def set_events(self, enable):
        check(library().vxhip_world_set_events(self._h, 1 if enable else 0))
        return self

    def set_sleep(self, config=None, **kw):
        """Device sleep update on (a SleepConfig, or keyword overrides of the
        defaults) or off (no arguments and enable=False via set_sleep(False))."""
        if config is False:
            check(library().vxhip_world_set_sleep(self._h, None))
            return self
        cfg = config if isinstance(config, SleepConfig) else SleepConfig(**kw)
        check(library().vxhip_world_set_sleep(self._h, C.byref(cfg)))
        return self

    def set_joints(self, joints, any_count=False):
        """Joint records (numpy array of the JOINT dtype, or raw 420-byte rows).
        any_count: solve sets of fewer than 64 on the device too
        (vxhip_world_set_joints_all)."""
        from .records import JOINT
        arr = np.ascontiguousarray(joints).view(np.uint8).reshape(-1, JOINT.itemsize).copy().view(JOINT).reshape(-1)
        fn = library().vxhip_world_set_joints_all if any_count else library().vxhip_world_set_joints
        check(fn(self._h, ptr(arr) if len(arr) else None, len(arr)))
        return self

    def joints(self):
        from .records import JOINT
        count = C.c_uint32(0)
        rc = library().vxhip_world_download_joints(self._h, None, 0, C.byref(count))
        if rc not in (OK, ERR_CAPACITY):
            check(rc)
        out = np.zeros(count.value, JOINT)
        check(library().vxhip_world_download_joints(self._h, ptr(out) if len(out) else None, len(out),
                                                    C.byref(count)))
        return out

    def joint_breaks(self):
        from .records import JOINT_BREAK
        count = C.c_uint32(0)
        rc = library().vxhip_world_download_joint_breaks(self._h, None, 0, C.byref(count))
        if rc not in (OK, ERR_CAPACITY):
            check(rc)
        out = np.zeros(count.value, JOINT_BREAK)
        check(library().vxhip_world_download_joint_breaks(self._h, ptr(out) if len(out) else None, len(out),
                                                          C.byref(count)))
        return out

    def set_multi_toi(self, **kw):
        cfg = MultiToiSettings(**kw)
        check(library().vxhip_world_set_multi_toi(self._h, C.byref(cfg)))
        return self

    def query_multi_toi(self, body, dt, cap=None):
        """World::queryMultiToi on the resident world: (hits sorted by (toi,
        body index), total hit count); `cap` limits the hits returned."""
        from .records import TOI_HIT
        n = C.c_uint32(0)
        out = np.zeros(self.n, TOI_HIT)  # at most one hit per other body
        check(library().vxhip_world_query_multi_toi(self._h, int(body), float(dt), ptr(out), len(out), C.byref(n)))
        want = n.value if cap is None else min(int(cap), n.value)
        return out[:want].copy(), n.value

    def shift_origin(self, offset):
        off = (C.c_float * 3)(*offset)
        check(library().vxhip_world_shift_origin(self._h, off))
        return self

    def save_state(self):
        """Host blob of the device state the next step reads (bytes)."""
        size = C.c_size_t(0)
        check(library().vxhip_world_state_size(self._h, C.byref(size)))
        buf = (C.c_uint8 * size.value)()
        check(library().vxhip_world_save_state(self._h, buf, size.value))
        return bytes(buf)

    def restore_state(self, blob):
        buf = (C.c_uint8 * len(blob)).from_buffer_copy(blob)
        check(library().vxhip_world_restore_state(self._h, buf, len(blob)))
        return self

    # -- differentiable form --------------------------------------------
    def record(self, enable=True):
        """Keep each step's frozen active set for download_record()."""
        check(library().vxhip_world_record(self._h, 1 if enable else 0))
        return self

    def download_record(self):
        """The last step's contacts in solve order (warm-start impulses) and,
        per contact, the previous step's solve position it started from
        (0xffffffff: none)."""
        count = C.c_uint32(0)
        check(library().vxhip_world_download_record(self._h, None, None, 0, C.byref(count)))
        cs = np.zeros(max(1, count.value), CONTACT)
        match = np.zeros(max(1, count.value), np.uint32)
        check(library().vxhip_world_download_record(self._h, ptr(cs), ptr(match), len(cs), C.byref(count)))
        return cs[:count.value], match[:count.value]

    def diff_forward(self, dt, steps, every, targets):
        """Rollout of `steps` steps keeping a checkpoint every `every`; per-world
        loss |centre of mass - target|^2 after it (targets: W x 3)."""
        targets = np.ascontiguousarray(targets, np.float32).reshape(-1)
        loss = np.zeros(len(targets) // 3, np.float32)
        check(library().vxhip_diff_forward(self._h, float(dt), int(steps), int(every), ptr(targets), ptr(loss)))
        return loss

    def diff_backward(self, loss_weight):
        """Gradients of sum_w loss_weight[w] * loss_w: (dv0 (n, 3), dw0 (n, 3),
        dfriction (n,), dmass (n,), stats)."""
        lw = np.ascontiguousarray(loss_weight, np.float32)
        gv = np.zeros((self.n, 3), np.float32)
        gw = np.zeros((self.n, 3), np.float32)
        gf = np.zeros(self.n, np.float32)
        gm = np.zeros(self.n, np.float32)
        st = DiffStats()
        check(library().vxhip_diff_backward(self._h, ptr(lw), ptr(gv), ptr(gw), ptr(gf), ptr(gm), C.byref(st)))
        return gv, gw, gf, gm, st.as_dict()

    def sleep_stats(self):
        st = SleepStats()
        check(library().vxhip_world_sleep_stats(self._h, C.byref(st)))
        return st.as_dict()

    def stats(self):
        s = Stats()
        check(library().vxhip_world_stats(self._h, C.byref(s)))
        return s.as_dict()

    # -- stage probes -----------------------------------------------------
    def probe_pairs(self, bodies, soup, dt, cap=None):
        bodies = np.ascontiguousarray(bodies)
        view = soup.view()
        cap = cap or max(1024, 64 * len(bodies))
        out = np.zeros(cap, np.uint64)
        cnt = C.c_uint32(0)
        code = library().vxhip_probe_pairs(self._h, ptr(bodies), len(bodies), C.byref(view), float(dt), ptr(out),
                                           cap, C.byref(cnt))
        if code == ERR_CAPACITY:
            return self.probe_pairs(bodies, soup, dt, cnt.value + 16)
        check(code)
        return out[:cnt.value]

    def find_contacts(self, bodies, soup, dt, cap=None):
        bodies = np.ascontiguousarray(bodies)
        view = soup.view()
        cap = cap or max(1024, 16 * len(bodies))
        out = np.zeros(cap, CONTACT)
        cnt = C.c_uint32(0)
        code = library().vxhip_find_contacts(self._h, ptr(bodies), len(bodies), C.byref(view), float(dt), ptr(out),
                                             cap, C.byref(cnt))
        if code == ERR_CAPACITY:
            return self.find_contacts(bodies, soup, dt, cnt.value + 16)
        check(code)
        return out[:cnt.value]

    def color(self, bodies, contacts):
        bodies = np.ascontiguousarray(bodies)
        contacts = np.ascontiguousarray(contacts)
        colors = np.zeros(max(1, len(contacts)), np.uint32)
        k = C.c_uint32(0)
        check(library().vxhip_probe_color(self._h, ptr(bodies), len(bodies), ptr(contacts), len(contacts),
                                          ptr(colors), C.byref(k)))
        return colors[:len(contacts)], k.value

    def radix_sort(self, keys, vals, bits):
        """(keys, vals) stably sorted by the low `bits` of the key on the device."""
        keys = np.ascontiguousarray(keys, np.uint64).copy()
        vals = np.ascontiguousarray(vals, np.uint32).copy()
        check(library().vxhip_probe_radix_sort(self._h, ptr(keys), ptr(vals), len(keys), int(bits)))
        return keys, vals

    def std_sort(self, keys, vals, seg_start=None):
        """(keys, vals) in the order libstdc++'s std::sort leaves them (per
        segment when seg_start is given), computed on the device."""
        keys = np.ascontiguousarray(keys, np.uint64).copy()
        vals = np.ascontiguousarray(vals, np.uint32).copy()
        seg = None if seg_start is None else np.ascontiguousarray(seg_start, np.uint32)
        check(library().vxhip_probe_std_sort(self._h, ptr(keys), ptr(vals), len(keys),
                                             ptr(seg) if seg is not None else None,
                                             0 if seg is None else len(seg) - 1))
        return keys, vals

    def warm_start_match(self, cur, prev):
        cur = np.ascontiguousarray(cur)
        prev = np.ascontiguousarray(prev)
        check(library().vxhip_probe_warm_start(self._h, ptr(cur), len(cur), ptr(prev), len(prev)))
        return cur

    def integrate(self, bodies, gravity, dt):
        g = (C.c_float * 3)(*gravity)
        check(library().vxhip_integrate(self._h, ptr(bodies), len(bodies), g, float(dt)))

    def solve_velocities(self, bodies, contacts, dt, warm_start, options=None):
        opts = options or default_options()
        check(library().vxhip_solve_velocities(self._h, ptr(bodies), len(bodies), ptr(contacts), len(contacts),
                                               float(dt), int(warm_start), C.byref(opts)))

    def advance_substeps(self, bodies, contacts, gravity, h, substeps, options=None, joints=None):
        """All substeps of one step in place on `bodies`, `contacts` (left in
        colour order) and `joints` (a JOINT array, 64 or more to run here);
        False when declined."""
        opts = options or default_options()
        g = (C.c_float * 3)(*gravity)
        adv = C.c_int(0)
        nj = 0 if joints is None else len(joints)
        if nj and (not joints.flags.c_contiguous or joints.nbytes != nj * 420):
            raise ValueError("joints must be a C-contiguous array of 420-byte joint records")
        check(library().vxhip_advance_substeps(self._h, ptr(bodies), len(bodies), ptr(contacts), len(contacts),
                                               ptr(joints) if nj else None, nj, g, float(h), int(substeps),
                                               C.byref(opts), C.byref(adv)))
        return bool(adv.value)

    def fetch_impulses(self, contacts):
        """The accumulated impulses of the last solve into `contacts` (in the
        order the warm-started solve left them); untouched when the count
        differs or nothing was solved since the last invalidation."""
        check(library().vxhip_fetch_impulses(self._h, ptr(contacts), len(contacts)))

    def invalidate_caches(self):
        check(library().vxhip_invalidate_caches(self._h))

    def last_velocity_iterations(self):
        return library().vxhip_last_velocity_iterations(self._h)
