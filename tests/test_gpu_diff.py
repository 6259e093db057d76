"""The differentiable form on the GPU: 1,024 worlds of 64 cubes, each pushed
by an initial impulse, loss |final centre of mass - target|^2 per world after
120 steps, gradients by the hand-written adjoint kernels behind
torch.autograd.Function with a checkpoint every 8 steps.

Checked against:
  * the double-precision CPU reference adjoint (oracle/vxo_diff.cpp) on the
    same frozen active set, recorded from the device (a sample of worlds,
    each stepped alone: batched worlds evolve bit for bit as alone);
  * double-precision central differences (h = 1e-3) of the reference forward
    on contact-stable parameters (perturbing them by +-h leaves every step's
    contact pairs and features unchanged on the device);
  * the tape replay reproducing the rollout bit for bit.
Tolerance: 1e-2 relative, as a vector norm over all of a world's parameters
and over each parameter group (v0, w0, friction, mass) that carries at least
5% of the largest group's gradient norm.  Needs an MI355X.
"""
import ctypes as C

import numpy as np
import pytest

from vxfix import oracle, ptr

pytestmark = pytest.mark.gpu

STEPS, EVERY = 120, 8


def _record(scene, bodies, steps):
    """A world stepped alone with recording on: per-step contacts, matches."""
    from velox_amd.device import Context
    ctx = Context()
    try:
        ctx.load(bodies, scene.soup, gravity=scene.gravity, substeps=scene.substeps)
        ctx.record(True)
        contacts, matches, keys = [], [], []
        for _ in range(steps):
            ctx.step(np.float32(scene.dt))
            cs, match = ctx.download_record()
            contacts.append(cs)
            matches.append(np.where(match == 0xffffffff, -1, match).astype(np.int32))
            keys.append(np.stack([cs["a"].astype(np.uint64), cs["b"].astype(np.uint64), cs["featureKey"]], 1))
        final = ctx.bodies()
    finally:
        ctx.close()
    return contacts, matches, keys, final


def _reference(scene, bodies, contacts, matches, target, params=None, grad=False):
    offsets = np.concatenate([[0], np.cumsum([len(c) for c in contacts])]).astype(np.uint32)
    cs = np.concatenate(contacts)
    match = np.concatenate(matches)
    g = (C.c_float * 3)(*scene.gravity)
    tgt = (C.c_double * 3)(*[float(x) for x in target])
    out = np.zeros(8 * len(bodies), np.float64) if grad else None
    com = np.zeros(3, np.float64)
    loss = oracle().vxo_diff_rollout(ptr(bodies), len(bodies), ptr(cs), ptr(match), ptr(offsets), len(contacts),
                                     np.float32(scene.dt), scene.substeps, 8, 3, g, tgt,
                                     ptr(params) if params is not None else None, ptr(out) if grad else None, ptr(com))
    return loss, (out.reshape(-1, 8) if grad else None), com


def _pack(v0, w0, fr, ms):
    return np.concatenate([v0, w0, fr[:, None], ms[:, None]], 1).astype(np.float64).reshape(-1)


def _rel(a, b):
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-12))


@pytest.fixture(scope="module")
def full_rollout():
    import torch
    from velox_amd.device import Scene
    from velox_amd.diff import Rollout, RolloutLoss
    scene = Scene("diff", 1024)
    ro = Rollout(scene, steps=STEPS, every=EVERY)
    v0, w0, fr, ms = (torch.tensor(a, device="cuda", requires_grad=True) for a in ro.initial())
    print(f"[gpu] diff rollout: {scene.worlds} worlds, {len(scene.bodies)} bodies, {STEPS} steps", flush=True)
    loss = RolloutLoss.apply(v0, w0, fr, ms, ro)
    loss.sum().backward()
    print(f"[gpu] diff stats: {ro.last_stats}", flush=True)
    grads = [t.grad.detach().cpu().numpy() for t in (v0, w0, fr, ms)]
    return scene, ro, loss.detach().cpu().numpy(), grads


def test_rollout_replay_is_bitwise_and_ccd_free(full_rollout):
    scene, ro, loss, grads = full_rollout
    st = ro.last_stats
    assert st["replayMismatch"] == 0, st
    assert st["ccdRescued"] == 0, st
    assert st["checkpoints"] == STEPS // EVERY and st["worlds"] == scene.worlds
    assert np.isfinite(loss).all() and all(np.isfinite(g).all() for g in grads)


@pytest.mark.parametrize("w", [0, 341, 1023])
def test_gradients_match_reference_adjoint(full_rollout, w):
    scene, ro, loss, (gv, gw, gf, gm) = full_rollout
    idx = np.nonzero(scene.world_of == w)[0]
    params = [a[idx] for a in ro.initial()]
    bodies = ro.bodies_with(*ro.initial())[idx].copy()
    contacts, matches, _, final = _record(scene, bodies, STEPS)
    ref_loss, ref, com = _reference(scene, bodies, contacts, matches, ro.targets[w], _pack(*params), grad=True)
    dyn = bodies["motionType"] == 2
    print(f"[gpu] world {w}: loss device {loss[w]:.6g} reference {ref_loss:.6g}, "
          f"{sum(len(c) for c in contacts)} contact-steps", flush=True)
    assert abs(loss[w] - ref_loss) <= 1e-3 * abs(ref_loss) + 1e-6
    groups = {"v0": (gv[idx][dyn], ref[dyn, 0:3]), "w0": (gw[idx][dyn], ref[dyn, 3:6]),
              "friction": (gf[idx][dyn], ref[dyn, 6]), "mass": (gm[idx][dyn], ref[dyn, 7])}
    errs = {}
    for name, (got, want) in groups.items():
        errs[name] = (_rel(got, want), float(np.linalg.norm(want)))
        print(f"[gpu]   {name}: |g| {errs[name][1]:.4g} rel err {errs[name][0]:.3g}", flush=True)
    got = np.concatenate([np.ravel(g) for g, _ in groups.values()])
    want = np.concatenate([np.ravel(r) for _, r in groups.values()])
    whole = _rel(got, want)
    print(f"[gpu]   all parameters: rel err {whole:.3g}", flush=True)
    assert whole < 1e-2
    # Each group on its own, where the group carries at least 5% of the
    # largest group's gradient (a near-zero group's relative error measures
    # FP32 rounding of the rollout, not the adjoint).
    top = max(n for _, n in errs.values())
    for name, (err, norm) in errs.items():
        if norm >= 0.05 * top:
            assert err < 1e-2, (name, err)


@pytest.mark.parametrize("w", [5, 700])
def test_gradients_match_central_differences(full_rollout, w):
    """FP64 central differences (h = 1e-3) of the reference forward on the
    frozen active set, for parameters whose +-h perturbation leaves the
    device's active set unchanged at every step."""
    scene, ro, loss, (gv, gw, gf, gm) = full_rollout
    idx = np.nonzero(scene.world_of == w)[0]
    base = [a[idx].copy() for a in ro.initial()]
    bodies = ro.bodies_with(*ro.initial())[idx].copy()
    contacts, matches, keys, _ = _record(scene, bodies, STEPS)
    dyn = np.nonzero(bodies["motionType"] == 2)[0]
    rng = np.random.default_rng(w)
    picks = [(0, i, 0) for i in rng.choice(dyn, 3, replace=False)] + [(1, rng.choice(dyn), 1)] + \
            [(2, rng.choice(dyn), None), (3, rng.choice(dyn), None)]
    device = {0: gv[idx], 1: gw[idx], 2: gf[idx], 3: gm[idx]}
    h = 1e-3
    got, want, stable = [], [], 0
    for slot, body, comp in picks:
        evals = []
        for sgn in (1.0, -1.0):
            p = [a.copy() for a in base]
            if comp is None:
                p[slot][body] += sgn * h
            else:
                p[slot][body, comp] += sgn * h
            # contact stability of the perturbed rollout on the device
            pert = ro.bodies_with(*[np.concatenate([a[:idx[0]], q, a[idx[-1] + 1:]]) for a, q in
                                    zip(ro.initial(), p)])[idx].copy()
            _, _, keys_p, _ = _record(scene, pert, STEPS)
            same = len(keys_p) == len(keys) and all(np.array_equal(a, b) for a, b in zip(keys_p, keys))
            evals.append((same, _reference(scene, bodies, contacts, matches, ro.targets[w], _pack(*p))[0]))
        if not (evals[0][0] and evals[1][0]):
            continue
        stable += 1
        fd = (evals[0][1] - evals[1][1]) / (2 * h)
        g = device[slot][body] if comp is None else device[slot][body, comp]
        got.append(g)
        want.append(fd)
    print(f"[gpu] world {w}: {stable} of {len(picks)} parameters contact-stable; device {np.round(got, 6)} "
          f"fd {np.round(want, 6)}", flush=True)
    assert stable >= 3, "too few contact-stable parameters to check"
    assert _rel(np.array(got), np.array(want)) < 1e-2
