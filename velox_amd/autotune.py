"""Autotuner of the device step's scheduling knobs (vxhip_tuning).

    python -m velox_amd.autotune [--workloads debris,batch] [--out velox_amd/tuning.json]

Runs on the GPU.  For each workload it takes snapshots of the resident world
at a few points of its trajectory (vxhip_world_save_state), then times short
windows restored from them under candidate knob settings: coordinate descent
from the library defaults, one knob at a time over its candidate values, two
rounds, each window timed twice and the faster kept.  Every candidate must
leave the world bit for bit where the defaults leave it (the knobs choose
schedules, never results); a difference aborts the run.

The table it writes (velox_amd/tuning.json) holds, per workload, the chosen
knobs, their measured ms per step against the defaults', and every trial.
Context.load_scene applies the entry of the scene's workload.

Knobs (include/vxhip.h, vxhip_tuning):
  batchThreads       workgroup size of the island-batch sweeps (rows per pass = half)
  persistentBatches  batch sweeps as one workgroup per CU taking batches from a queue
  longestFirst       batches handed out longest first (pass count)
  colorLaunchRows    big-island colours wider than this get a launch each
  islandLimit        islands above this many contacts go colour by colour over the device
  sideStream         island batches beside the big-island chain on a second stream
  graphSubsteps      substeps 2..S replayed from a captured hipGraph
  jointPackMax       joint packs above this size take a launch per colour (not
                     tuned here: the workloads' joint components are small)
"""
import argparse
import json
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(HERE))

from velox_amd.device import Context, Scene  # noqa: E402
from velox_amd.tuning import TABLE  # noqa: E402

CANDIDATES = {
    "batchThreads": [256, 512],
    "persistentBatches": [True, False],
    "longestFirst": [True, False],
    "colorLaunchRows": [512, 1024, 2048],
    "islandLimit": [2048, 4096, 8192],
    "sideStream": [True, False],
    "graphSubsteps": [False, True],
}
DEFAULTS = {"batchThreads": 256, "persistentBatches": True, "longestFirst": True, "colorLaunchRows": 1024,
            "islandLimit": 4096, "sideStream": True, "graphSubsteps": False}

# Workloads: scene generator and parameter, the steps at which snapshots are
# taken, steps timed per window and each window's weight (the share of the
# full run's steps it stands for).
WORKLOADS = {
    "debris": {"scene": ("debris", 50000), "windows": [(30, 5, 100.0 / 120.0), (105, 5, 20.0 / 120.0)],
               "about": "50,000-body mixed debris; windows after steps 30 (free fall and landing) and 105 "
                        "(the merged pile with its big island), weighted 100:20 as in steps 1-120"},
    "batch": {"scene": ("batch", 4096), "windows": [(20, 6, 1.0)],
              "about": "4,096 batched worlds (8-512 bodies each); window after step 20"},
}


def state_bits(ctx):
    b = ctx.bodies()
    return np.ascontiguousarray(b).view(np.uint8).tobytes()


class Bench:
    def __init__(self, name):
        spec = WORKLOADS[name]
        self.spec = spec
        self.scene = Scene(*spec["scene"])
        self.ctx = Context()
        self.ctx.set_tuning(**DEFAULTS)
        self.ctx.load_scene(self.scene, tuned=False)
        self.ctx.set_tuning(**DEFAULTS)
        self.snaps = []
        done = 0
        for at, _, _ in spec["windows"]:
            self.ctx.step(self.scene.dt, at - done)
            done = at
            self.ctx.sync()
            self.snaps.append(self.ctx.save_state())
        self.expect = None

    def run(self, knobs):
        """Weighted ms per step of `knobs` over the windows (best of two), and
        the end states of the windows."""
        self.ctx.set_tuning(**knobs)
        total, ends = 0.0, []
        for (at, steps, weight), snap in zip(self.spec["windows"], self.snaps):
            best = None
            for _ in range(2):
                self.ctx.restore_state(snap)
                self.ctx.step(self.scene.dt, 1)  # untimed: buffers sized for this window
                self.ctx.sync()
                self.ctx.restore_state(snap)
                self.ctx.sync()
                t0 = time.perf_counter()
                self.ctx.step(self.scene.dt, steps)
                self.ctx.sync()
                dt = (time.perf_counter() - t0) * 1e3 / steps
                best = dt if best is None else min(best, dt)
            ends.append(state_bits(self.ctx))
            total += weight * best
        return total / sum(w for _, _, w in self.spec["windows"]), ends


def tune(name, rounds=2, log=print):
    b = Bench(name)
    base_ms, b.expect = b.run(DEFAULTS)
    log(f"[tune] {name}: defaults {base_ms:.3f} ms/step")
    trials = [{"knobs": dict(DEFAULTS), "ms_per_step": round(base_ms, 4)}]
    best, best_ms = dict(DEFAULTS), base_ms
    for rnd in range(rounds):
        changed = False
        for knob, values in CANDIDATES.items():
            for v in values:
                if v == best[knob]:
                    continue
                cand = dict(best, **{knob: v})
                ms, ends = b.run(cand)
                if ends != b.expect:
                    raise RuntimeError(f"{name}: knobs {cand} changed the result")
                trials.append({"knobs": cand, "ms_per_step": round(ms, 4)})
                log(f"[tune] {name} round {rnd} {knob}={v}: {ms:.3f} ms/step (best {best_ms:.3f})")
                if ms < best_ms * 0.99:
                    best, best_ms, changed = cand, ms, True
        if not changed:
            break
    # The chosen setting again, against the defaults, back to back.
    final_ms, ends = b.run(best)
    if ends != b.expect:
        raise RuntimeError(f"{name}: knobs {best} changed the result")
    default_ms, _ = b.run(DEFAULTS)
    b.ctx.close()
    return {"knobs": best, "ms_per_step": round(final_ms, 4), "default_ms_per_step": round(default_ms, 4),
            "about": WORKLOADS[name]["about"], "trials": trials}


def main():
    ap = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawDescriptionHelpFormatter)
    ap.add_argument("--workloads", default="debris,batch")
    ap.add_argument("--out", default=TABLE)
    ap.add_argument("--rounds", type=int, default=2)
    args = ap.parse_args()
    import torch
    table = {"device": torch.cuda.get_device_name(0),
             "command": "python -m velox_amd.autotune " + " ".join(sys.argv[1:]), "workloads": {}}
    if os.path.exists(args.out):
        with open(args.out) as f:
            table["workloads"] = json.load(f).get("workloads", {})
    for name in args.workloads.split(","):
        table["workloads"][name] = tune(name, args.rounds, log=lambda s: print(s, flush=True))
        with open(args.out, "w") as f:
            json.dump(table, f, indent=1)
    print(json.dumps({k: {"knobs": v["knobs"], "ms_per_step": v["ms_per_step"],
                          "default_ms_per_step": v["default_ms_per_step"]}
                      for k, v in table["workloads"].items()}))


if __name__ == "__main__":
    main()
