#!/usr/bin/env python3
"""Throughput of the HIP device step on one MI355X.

Workload: the 50,000-body mixed-debris scene (30% spheres, 30% boxes, 20%
capsules, 20% 24-point hulls above a plane, FP32, dt = 1/60, 4 substeps,
8 velocity iterations).  A step is one full device step of the resident
world (integration, broad phase, narrow phase, warm-start matching, greedy
colouring, coloured TGS sweeps over 4 substeps, CCD net, positional
correction).  Inputs are resident in HBM before the timed region.

Window: the run's steps 1-120.  The configuration names 500 steps, but the
reference itself runs away past step ~130 on this scene (thin capsules spun
about their axis inflate their swept boxes until they pair with everything:
the reference's own contacts go from 0.1 M at step 131 to 1.2 M at step 141,
profiles/debris50k_reference_runaway.md), so 1-120 is the physically valid
part, and it holds the heavy regime (steps 100-120: one island of ~58,000
contacts, 100-230 colours).  The K timed steps (--steps) are spread evenly
over the window (every 120/K-th step; K = 120 times all of them), each
bracketed by a device synchronisation; the steps between them run untimed.
The reference CPU backend is timed on the same sampled steps.

Prints one JSON line: body-steps/s over the sampled steps, the contiguous
window's rate, the roofline of the event-timed kernel with the largest device
time (algorithmic bytes per launch over its event-timed duration), the
reference CPU baseline (setWorkerCount(0), median of 3 passes, plus a
1-thread figure), the batched-worlds comparison and the differentiable
form's gradient throughput.
"""
import argparse
import json
import os
import platform
import subprocess
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
WINDOW_END = 120
# Algorithmic bytes per work unit of each event-timed kernel (the step's byte
# model): a solve contact visit reads 96 B of rows, writes 24 B of impulses,
# and per body reads 28 B pose + 32 B mass and reads and writes 24 B of
# velocity; a candidate pair costs 272 B in the narrow phase (two 132-B
# geometry records and the 8-B key); a CCD probe 224 B; pair emission 40 B
# per body.
SOLVE_BYTES_PER_VISIT = 96 + 24 + 2 * (28 + 32 + 24 + 24)
BYTES_PER_UNIT = {"kBatchSweeps<0>": SOLVE_BYTES_PER_VISIT, "kBatchSweeps<2>": SOLVE_BYTES_PER_VISIT,
                  "kIslandSweeps": SOLVE_BYTES_PER_VISIT, "kSweepColorPair/Run/TailPair": SOLVE_BYTES_PER_VISIT,
                  "kNarrow": 272, "kCcdCheck": 224, "kCcdDataflow": 224, "kGridPairs": 40}


def host_info():
    model = None
    try:
        out = subprocess.run(["lscpu"], capture_output=True, text=True, timeout=10).stdout
        for line in out.splitlines():
            if line.startswith("Model name:"):
                model = line.split(":", 1)[1].strip()
    except (OSError, subprocess.SubprocessError):
        pass
    return {"nproc": os.cpu_count(), "cpu_model": model or platform.processor()}


def measured_traffic(kernel, window):
    """HBM bytes per launch of `kernel` measured by the committed rocprofv3
    FETCH_SIZE/WRITE_SIZE passes over this window (profiles/traffic.json), or
    None when there is no such measurement."""
    path = os.path.join(ROOT, "profiles", "traffic.json")
    if not os.path.exists(path):
        return None
    with open(path) as f:
        entry = json.load(f).get(f"{kernel}|{window}")
    return None if entry is None else entry["bytes_per_launch"]


def sampled_steps(k):
    k = max(1, min(int(k), WINDOW_END))
    stride = WINDOW_END // k
    return [stride * (i + 1) for i in range(k)], stride


def run_trajectory(ctx, scene, timed):
    """Steps 1..WINDOW_END from the scene's start; the steps in `timed` are
    each bracketed by a synchronisation and timed.  Returns their times."""
    ctx.load_scene(scene)
    ctx.sync()
    times = []
    dt = scene.dt
    todo = set(timed)
    for t in range(1, WINDOW_END + 1):
        if t in todo:
            ctx.sync()
            t0 = time.perf_counter()
            ctx.step(dt, 1)
            ctx.sync()
            times.append(time.perf_counter() - t0)
        else:
            ctx.step(dt, 1)
        if t % 20 == 0:
            st = ctx.stats()
            print(f"[bench] step {t}/{WINDOW_END} contacts={st['contacts']} colors={st['colors']} "
                  f"big={st['bigIslandContacts']}", file=sys.stderr, flush=True)
    return times


def profile_pass(ctx, scene):
    """The window again with per-launch HIP events on the library's stream."""
    from velox_amd.device import Timings
    ctx.load_scene(scene)
    ctx.profile(True)
    ctx.step(scene.dt, WINDOW_END)
    ctx.sync()
    t = ctx.timings()
    ctx.profile(False)
    kernels = {}
    for kind, name in enumerate(Timings.KERNELS):
        n = t["kernelLaunches"][kind]
        if n:
            kernels[name] = {"ms_per_step": round(t["kernelMs"][kind] / max(1, t["steps"]), 4), "launches": n,
                             "avg_launch_us": round(t["kernelMs"][kind] * 1e3 / n, 3),
                             "units": t["kernelVisits"][kind]}
    stages = {k[:-2]: round(v / max(1, t["steps"]), 4) for k, v in t.items() if k.endswith("Ms") and k != "kernelMs"}
    return kernels, stages


def roofline_of(kernels, window):
    """Roofline of the event-timed kernel with the largest device time."""
    name = max(kernels, key=lambda k: kernels[k]["ms_per_step"])
    k = kernels[name]
    bytes_per_launch = BYTES_PER_UNIT[name] * k["units"] / k["launches"]
    achieved = bytes_per_launch / (k["avg_launch_us"] * 1e-6) / 1e9
    return {"kernel": name, "bound": "hbm", "achieved": round(achieved, 3), "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": round(achieved / HBM_PEAK_GBS, 6), "traffic": measured_traffic(name, window),
            "avg_launch_us": k["avg_launch_us"], "launches": k["launches"],
            "bytes_per_launch": round(bytes_per_launch, 1), "bytes_per_unit": BYTES_PER_UNIT[name]}


def cpu_baseline(bodies, timed, stride, n, gpu_rate):
    exe = os.path.join(ROOT, "oracle", "_ref", "ref_bench")
    if not os.path.exists(exe):
        return None

    def run(end, workers, repeats):
        out = subprocess.run([exe, "sample", f"debris{bodies}", str(end), str(stride), str(workers), str(repeats)],
                             stdout=subprocess.PIPE, text=True, timeout=1500)
        if out.returncode != 0:
            return None
        return json.loads(out.stdout.strip().splitlines()[-1])

    full = run(WINDOW_END, 0, 3)
    if not full:
        return None
    one_end = min(WINDOW_END, 5 * stride)
    one = run(one_end, 1, 1)
    info = host_info()
    return {"value": round(full["median"], 1), "unit": "body-steps/s", "cores": full["workers"], "kind": "reference",
            "sample": f"reference World(BackendType::Cpu), relaxed -O3 build, setWorkerCount(0) -> "
                      f"{full['workers']} workers (hardware_concurrency {full['hardware_concurrency']}), the same "
                      f"{len(timed)} sampled steps (every {stride}th of steps 1-{WINDOW_END}) of its own run of the "
                      f"{n}-body debris scene, median of 3 passes (one along the trajectory, two restoring World "
                      f"snapshots before each sampled step)",
            "passes": [round(r, 1) for r in full["rates"]],
            "single_thread": {"value": round(one["median"], 1) if one else None,
                              "sample": f"setWorkerCount(1), the sampled steps among steps 1-{one_end}"},
            "nproc": info["nproc"], "cpu_model": info["cpu_model"],
            "gpu_same_sample": round(gpu_rate, 1), "ratio": round(gpu_rate / full["median"], 2)}


def batched_leg(worlds, steps, sample, plain_steps):
    """World-steps/s of one batched device step over `worlds` independent
    worlds against the same worlds stepped one after another, each alone (a
    uniform stride sample of them; every world's cost is its own, so the
    sample's rate estimates the whole set's)."""
    from velox_amd.device import Context, Scene
    scene = Scene("batch", worlds)
    dt = scene.dt
    ctx = Context()
    ctx.load_scene(scene)
    ctx.step(dt, 1)  # warm-up (buffer growth)
    ctx.sync()
    ctx.load_scene(scene)
    ctx.sync()
    t0 = time.perf_counter()
    done = 0
    while done < steps:
        k = min(50, steps - done)
        ctx.step(dt, k)
        ctx.sync()
        done += k
        print(f"[bench] batched step {done}/{steps} t={time.perf_counter() - t0:.2f}s", file=sys.stderr, flush=True)
    batched_s = time.perf_counter() - t0
    st = ctx.stats()
    ctx.close()
    stride = max(1, worlds // max(1, sample))
    chosen = list(range(0, worlds, stride))[:sample]
    plain = Context()
    plain_s = 0.0
    for w in chosen:
        sub = scene.world(w)
        plain.load_scene(sub)
        plain.step(dt, 1)
        plain.sync()
        plain.load_scene(sub)
        plain.sync()
        t0 = time.perf_counter()
        for _ in range(plain_steps):
            plain.step(dt, 1)
            plain.sync()
        plain_s += time.perf_counter() - t0
    plain.close()
    print(f"[bench] plain: {len(chosen)} worlds x {plain_steps} steps in {plain_s:.3f}s", file=sys.stderr, flush=True)
    batched_rate = worlds * steps / batched_s
    plain_rate = len(chosen) * plain_steps / plain_s
    return {"metric": "batched vs plain world-steps/s, same GPU", "worlds": worlds, "bodies": int(len(scene.bodies)),
            "batched_world_steps_per_s": round(batched_rate, 1), "plain_world_steps_per_s": round(plain_rate, 1),
            "ratio": round(batched_rate / plain_rate, 2), "batched_steps": steps,
            "batched_ms_per_step": round(batched_s * 1e3 / steps, 3), "contacts_last_step": st["contacts"],
            "plain": f"{len(chosen)} of {worlds} worlds (every {stride}th), each loaded alone and stepped through its "
                     f"first {plain_steps} steps with a sync per world-step (timed), after an untimed warm-up step",
            "joints": int(len(scene.joints)),
            "workload": "log-uniform 8..512 bodies per world (box towers, sphere piles, ball-jointed capsule chains) "
                        "on one plane each, dt=1/60, 4 substeps, 8 iterations"}


def diff_leg(worlds, steps, every):
    """Gradient throughput of the differentiable form: rollout + adjoint of
    `worlds` worlds of 64 cubes over `steps` steps, checkpoint every `every`."""
    import numpy as np
    from velox_amd.device import Scene
    from velox_amd.diff import Rollout
    scene = Scene("diff", worlds)
    ro = Rollout(scene, steps=steps, every=every)
    p = ro.initial()
    ro.forward(*p)  # warm-up (buffer growth)
    ro.backward(np.ones(worlds, np.float32))
    t0 = time.perf_counter()
    loss = ro.forward(*p)
    t1 = time.perf_counter()
    ro.backward(np.ones(worlds, np.float32))
    t2 = time.perf_counter()
    print(f"[bench] diff: {worlds} worlds x {steps} steps, forward {t1 - t0:.3f}s backward {t2 - t1:.3f}s",
          file=sys.stderr, flush=True)
    return {"metric": "gradient world-steps/s (rollout + adjoint)", "worlds": worlds, "bodies": int(len(scene.bodies)),
            "steps": steps, "checkpoint_every": every, "forward_s": round(t1 - t0, 4), "backward_s": round(t2 - t1, 4),
            "gradient_world_steps_per_s": round(worlds * steps / (t2 - t0), 1),
            "mean_loss": round(float(loss.mean()), 6), "stats": ro.last_stats,
            "workload": "64 cubes per world pushed by an initial impulse, loss |final centre of mass - target|^2, "
                        "gradients w.r.t. initial linear and angular velocity, friction and mass"}


def main():
    ap = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawDescriptionHelpFormatter)
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=WINDOW_END, help="timed steps, spread over steps 1-120")
    ap.add_argument("--warmup", type=int, default=1, help="untimed steps of a separate run before the timed one")
    ap.add_argument("--bodies", type=int, default=50000)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-batched", action="store_true", help="skip the batched-worlds leg")
    ap.add_argument("--no-diff", action="store_true", help="skip the differentiable leg")
    ap.add_argument("--no-profile", action="store_true")
    ap.add_argument("--batched-only", action="store_true", help="run only the batched-worlds leg (profiling)")
    ap.add_argument("--worlds", type=int, default=4096, help="batched leg: number of worlds")
    ap.add_argument("--batched-steps", type=int, default=300, help="batched leg: timed steps of the batched world")
    ap.add_argument("--plain-sample", type=int, default=512,
                    help="batched leg: worlds stepped one by one (uniform stride sample)")
    ap.add_argument("--plain-steps", type=int, default=3, help="batched leg: timed steps per plain world")
    ap.add_argument("--diff-worlds", type=int, default=1024)
    ap.add_argument("--diff-steps", type=int, default=120)
    ap.add_argument("--calibrate-copy", type=int, default=0, metavar="MIB",
                    help="run one plain device copy of this size first (HBM counter calibration under rocprofv3)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py runs on one GPU: the device step has no multi-GPU path (asked for --gpus {args.gpus})",
              file=sys.stderr)
        sys.exit(2)

    if args.calibrate_copy:
        import torch
        src = torch.ones(args.calibrate_copy << 18, dtype=torch.float32, device="cuda")
        dst = torch.empty_like(src)
        dst.copy_(src)
        torch.cuda.synchronize()
        del src, dst

    from velox_amd.device import Context, Scene
    if args.batched_only:
        print(json.dumps({"batched": batched_leg(args.worlds, args.batched_steps, args.plain_sample,
                                                 args.plain_steps)}))
        return
    scene = Scene("debris", args.bodies)
    n = len(scene.bodies)
    ctx = Context()
    if args.warmup:
        ctx.load_scene(scene)
        ctx.step(scene.dt, args.warmup)
        ctx.sync()
    timed, stride = sampled_steps(args.steps)
    times = run_trajectory(ctx, scene, timed)
    seconds = sum(times)
    k = len(timed)
    value = n * k / seconds
    last = ctx.stats()
    # The contiguous window, for reference.
    ctx.load_scene(scene)
    ctx.sync()
    t0 = time.perf_counter()
    ctx.step(scene.dt, WINDOW_END)
    ctx.sync()
    window_s = time.perf_counter() - t0
    window_key = f"debris{args.bodies}|steps1-{WINDOW_END}"

    roofline = kernels = stages = None
    if not args.no_profile:
        kernels, stages = profile_pass(ctx, scene)
        roofline = roofline_of(kernels, window_key)
    ctx.close()

    cpu = None if args.no_cpu else cpu_baseline(args.bodies, timed, stride, n, value)
    batched = None if args.no_batched else batched_leg(args.worlds, args.batched_steps, args.plain_sample,
                                                       args.plain_steps)
    diff = None if args.no_diff else diff_leg(args.diff_worlds, args.diff_steps, 8)

    line = {
        "metric": "body-steps/s @50k mixed bodies",
        "value": round(value, 1),
        "unit": "body-steps/s",
        "n_gpus": 1,
        "steps": k,
        "warmup": args.warmup,
        "ms_per_step": round(seconds * 1e3 / k, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": f"mixed debris, {args.bodies} bodies + plane, dt=1/60, 4 substeps, 8 iterations; "
                               f"{k} timed steps spread over steps 1-{WINDOW_END} (every {stride}th), each "
                               f"synchronised, the steps between them untimed",
                   "bodies": n, "window": f"steps 1-{WINDOW_END}", "sampled_steps": timed,
                   "contacts_last_step": last["contacts"], "colors_last_step": last["colors"],
                   "parallelism": "single GPU"},
        "window_body_steps_per_s": round(n * WINDOW_END / window_s, 1),
        "roofline": roofline,
        "cpu_baseline": cpu,
        "batched": batched,
        "differentiable": diff,
    }
    if kernels:
        line["stage_ms_per_step"] = stages
        line["kernels"] = kernels
    print(json.dumps(line))


if __name__ == "__main__":
    main()
