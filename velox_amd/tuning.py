"""The autotuner's table: scheduling knobs of the device step per workload.

velox_amd/autotune.py measures the knobs of include/vxhip.h's vxhip_tuning on
the GPU for each workload and writes velox_amd/tuning.json; Context.load_scene
applies the entry of the scene's workload.  No knob changes a result (the
autotuner checks that every candidate steps the world bit for bit the same),
so a missing or stale table only costs time.
"""
import json
import os

TABLE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "tuning.json")


def workload_of(scene):
    """The table key of a synthetic scene: its generator name ("debris",
    "batch", ...), or "default"."""
    return getattr(scene, "name", None) or "default"


def load_table(path=TABLE):
    if not os.path.exists(path):
        return {}
    with open(path) as f:
        return json.load(f).get("workloads", {})


def knobs_for(scene, path=TABLE):
    """The tuned knobs for `scene` (empty: the library defaults)."""
    table = load_table(path)
    entry = table.get(workload_of(scene)) or table.get("default") or {}
    return dict(entry.get("knobs", {}))
