// Conservative-advancement safety net on the device (world.cpp:4227-4357).
//
// The reference walks the unique contact pairs in key order on the host, and
// a rescue rewinds bodies that later pairs read.  kCcdCheck evaluates every
// pair in parallel against the end-of-substeps state; kCcdDataflow then
// reproduces the reference order exactly, revisiting a pair only when it was
// flagged or one of its bodies was moved by an earlier rescue (any other pair
// sees the same state the parallel check saw).
#pragma once
#include "vx_kernels.hpp"

namespace vxk {

struct Gap { float gap; V3 normal; };

// `coop`: the whole wavefront evaluates this same query (hull supports are
// then shared across the lanes; see hullSupportIndexWave).
__device__ __attribute__((always_inline)) inline Gap gapSimple(const vx_body& ta, const vx_body& tb, const Soup& soup,
                                                               float search, bool coop = false) {
    if (tb.shape == VX_SHAPE_PLANE) {
        Convex c = makeConvex(ta, soup);
        if (coop) c.makeCoop();
        V3 pn = v3(tb.planeNormal);
        V3 deep = c.support(-pn);
        return {dot(pn, deep) - tb.planeOffset - c.radius, pn};
    }
    if (tb.shape == VX_SHAPE_MESH) {
        Gap best{1e30f, {0, 1, 0}};
        meshGapProbe1(ta, tb, soup, search, best.gap, best.normal);
        return best;
    }
    Convex ca = makeConvex(ta, soup), cb = makeConvex(tb, soup);
    if (coop) {
        ca.makeCoop();
        cb.makeCoop();
    }
    GjkResult r = gjkDistance(ca, cb);
    return {r.distance, r.normal};
}

// world.cpp:2916-2958 (compound children expanded without recursion): the
// child pairs are visited in the reference's nested order (children of a,
// then of b) and the first smallest gap is kept, which is what the nested
// strict-minimum loops keep; a pair of simple shapes returns its gap as it
// is.  One gapSimple call site, so the query inlines as one copy: a call per
// query would save and restore the callee's registers through scratch memory
// on every conservative-advancement iteration.
__device__ __attribute__((always_inline)) inline Gap gapAtInline(const vx_body& a0, V3 pa, Q4 qa, const vx_body& b0,
                                                                 V3 pb, Q4 qb, const Soup& soup, float search,
                                                                 bool coop = false) {
    vx_body ta = a0; ta.position = raw(pa); ta.orientation = raw(qa);
    vx_body tb = b0; tb.position = raw(pb); tb.orientation = raw(qb);
    const bool ca = ta.shape == VX_SHAPE_COMPOUND, cb = tb.shape == VX_SHAPE_COMPOUND;
    const uint32_t na = ca ? ta.compoundCount : 1u, nb = cb ? tb.compoundCount : 1u;
    Gap best{1e30f, {0, 1, 0}};
    for (uint32_t i = 0; i < na; ++i)
        for (uint32_t k = 0; k < nb; ++k) {
            vx_body x = ta, y = tb;
            if (ca) x = childBody(ta, soup.children[ta.compoundFirst + i]);
            if (cb) y = childBody(tb, soup.children[tb.compoundFirst + k]);
            const Gap g = gapSimple(x, y, soup, search, coop);
            if (!ca && !cb) return g;
            if (g.gap < best.gap) best = g;
        }
    return best;
}

__device__ inline Gap gapAt(const vx_body& a0, V3 pa, Q4 qa, const vx_body& b0, V3 pb, Q4 qb, const Soup& soup,
                            float search, bool coop = false) {
    return gapAtInline(a0, pa, qa, b0, pb, qb, soup, search, coop);
}

__device__ inline Q4 orientationAt(const vx_body& body, Q4 start, V3 L, float elapsed) {
    vx_body s = body;
    s.orientation = raw(start);
    s.angularVelocity = raw(invInertiaMul(s, L));
    advanceOrientation(s, elapsed);
    return q4(s.orientation);
}

__device__ inline void ccdBodies(const vx_body* bodies, uint64_t key, int nb, uint32_t& i, uint32_t& j) {
    i = (uint32_t)(key >> nb);
    j = (uint32_t)(key & ((1ull << nb) - 1));
    if (!isDynamic(bodies[i]) && isDynamic(bodies[j])) { uint32_t t = i; i = j; j = t; }
}

__device__ inline bool ccdEligible(const vx_body& a, const vx_body& b) {
    if (isSensor(a) || isSensor(b) || !isDynamic(a)) return false;
    return !(!a.ccdTuning.enableContinuous || a.ccdTuning.quality == VX_QUALITY_LOW || isLocked(a) ||
             maxPointSpeed(a) < a.ccdTuning.minVelocityForCCD);
}

__device__ inline float ccdSearch(const vx_body& a, float dt) {
    return maxPointSpeed(a) * dt + a.radius + length(v3(a.halfExtents)) + a.capsuleHalfHeight + 0.1f;
}

// Flags pairs (that produced contacts) whose end state penetrates deeper than
// 8 mm and records the first flagged index.
__global__ __launch_bounds__(kCollideThreads) void kCcdCheck(const vx_body* __restrict__ bodies, Soup soup, float dt,
                                                      const uint64_t* __restrict__ pairs,
                                                      const uint32_t* __restrict__ counts, uint32_t np, int nb,
                                                      const uint32_t* __restrict__ order,
                                                      uint8_t* __restrict__ flag, float4* __restrict__ endGap,
                                                      uint32_t* __restrict__ flagged,
                                                      uint32_t* __restrict__ firstFlagged,
                                                      const uint8_t* __restrict__ toiDone) {
    const uint32_t q = blockIdx.x * kCollideThreads + threadIdx.x;
    if (q >= np) return;
    const uint32_t p = order[q];  // pairs grouped by shape class
    uint8_t f = 0;
    if (counts[p]) {
        uint32_t i, j;
        ccdBodies(bodies, pairs[p], nb, i, j);
        const vx_body a = bodies[i], b = bodies[j];
        // Bodies the multi-TOI scheduler replayed are left alone (world.cpp:4268-4269).
        if (ccdEligible(a, b) && !(toiDone && (toiDone[i] || toiDone[j]))) {
            Gap end = gapAtInline(a, v3(a.position), q4(a.orientation), b, v3(b.position), q4(b.orientation), soup,
                            ccdSearch(a, dt));
            if (end.gap < -8e-3f) {
                f = 1;
                endGap[p] = make_float4(end.normal.x, end.normal.y, end.normal.z, end.gap);
            }
        }
    }
    flag[p] = f;
    if (f) {
        atomicAdd(flagged, 1u);
        atomicMin(firstFlagged, p);
    }
}

// The reference's per-pair rescue (world.cpp:4276-4355) on one lane.
// `known` carries the end gap the parallel check measured when neither body
// has moved since (same state, same result).
// With `coop` the whole wavefront runs the same rescue (every lane computes
// the same values) and only the `writer` lane stores the bodies.
__device__ __attribute__((noinline)) bool ccdRescue(vx_body* bodies, const vx_vec3* prevPos, const vx_quat* prevQ,
                                                    uint32_t i, uint32_t j, const Soup& soup, float dt,
                                                    const float4* known, bool coop = false, bool writer = true) {
    vx_body a = bodies[i], b = bodies[j];
    if (!ccdEligible(a, b)) return false;
    const float search = ccdSearch(a, dt);
    V3 pa0 = v3(prevPos[i]), da = v3(a.position) - pa0;
    Q4 qa0 = q4(prevQ[i]);
    V3 pb0 = isStatic(b) ? v3(b.position) : v3(prevPos[j]);
    V3 db = v3(b.position) - pb0;
    Q4 qb0 = isStatic(b) ? q4(b.orientation) : q4(prevQ[j]);
    float bound = maxPointSpeed(a) * dt + (isStatic(b) ? 0.0f : maxPointSpeed(b) * dt);
    V3 La = angularMomentum(a);
    V3 Lb = isDynamic(b) ? angularMomentum(b) : V3{};
    // Pass -1 measures the end-of-step gap (when the parallel check's value
    // cannot be reused): a pair at most 8 mm deep, or whose bodies barely
    // move, is left alone.  Passes 0..23 advance conservatively from the
    // step's start.  One gap query call site for all passes.
    Gap end{};
    if (known) end = {known->w, {known->x, known->y, known->z}};
    float t = 0.0f;
    V3 n{};
    for (int it = known ? 0 : -1;; ++it) {
        if (it == 0) {
            if (end.gap >= -8e-3f || bound < 1e-9f) return false;
            n = end.normal;
        }
        if (it >= 0 && !(it < 24 && t < 1.0f)) break;
        V3 pA, pB;
        Q4 qA, qB;
        if (it < 0) {
            pA = v3(a.position); qA = q4(a.orientation);
            pB = v3(b.position); qB = q4(b.orientation);
        } else {
            qA = orientationAt(a, qa0, La, dt * t);
            qB = isStatic(b) ? qb0 : orientationAt(b, qb0, Lb, dt * t);
            pA = pa0 + da * t;
            pB = pb0 + db * t;
        }
        const Gap g = gapAtInline(a, pA, qA, b, pB, qB, soup, search, coop);
        if (it < 0) {
            end = g;
            continue;
        }
        n = g.normal;
        if (g.gap < 1e-3f) break;
        t = fmn(1.0f, t + g.gap / bound);
    }
    if (t >= 1.0f) return false;
    a.position = raw(pa0 + da * t);
    a.orientation = raw(qa0);
    a.angularVelocity = raw(invInertiaMul(a, La));
    advanceOrientation(a, dt * t);
    if (isDynamic(b)) {
        b.position = raw(pb0 + db * t);
        b.orientation = raw(qb0);
        b.angularVelocity = raw(invInertiaMul(b, Lb));
        advanceOrientation(b, dt * t);
    }
    float vn = dot(v3(a.velocity) - v3(b.velocity), n);
    float ims = solverInvMass(a) + solverInvMass(b);
    if (vn < 0.0f && ims > 0.0f) {
        float rest = vn < -1.0f ? combine(a.restitution, b.restitution, a.restitutionCombine, b.restitutionCombine) : 0.0f;
        float imp = -(1.0f + rest) * vn / ims;
        a.velocity = raw(v3(a.velocity) + n * (imp * solverInvMass(a)));
        if (isDynamic(b)) b.velocity = raw(v3(b.velocity) - n * (imp * solverInvMass(b)));
    }
    float remaining = dt * (1.0f - t);
    advanceTransform(a, remaining);
    if (isDynamic(b)) advanceTransform(b, remaining);
    if (!writer) return true;
    bodies[i].position = a.position; bodies[i].orientation = a.orientation;
    bodies[i].velocity = a.velocity; bodies[i].angularVelocity = a.angularVelocity;
    if (isDynamic(b)) {
        bodies[j].position = b.position; bodies[j].orientation = b.orientation;
        bodies[j].velocity = b.velocity; bodies[j].angularVelocity = b.angularVelocity;
    }
    return true;
}

constexpr int kCcdThreads = 1024;

// Pairs that can see a rewound body.  A rescue only happens at a flagged
// pair or at a pair one of whose dynamic bodies an earlier rescue moved, so
// every pair that can do anything is flagged or follows (in pair order) such
// a pair on a shared dynamic body.  Each relevant pair lowers firstRel of its
// dynamic bodies to its index, and a pair behind a body's firstRel becomes
// relevant, until nothing changes.  A pair that becomes relevant lowers its
// bodies' firstRel at once, within the same round: firstRel only ever holds
// indices of relevant pairs, so any pair it puts behind one is in the
// closure, and a round that adds nothing has seen every earlier round's
// marks.  Every other pair sees exactly the state the parallel check saw.
__global__ __launch_bounds__(kThreads) void kCcdFlagsToRel(const vx_body* __restrict__ bodies,
                                                           const uint64_t* __restrict__ pairs,
                                                           const uint8_t* __restrict__ flag, uint32_t np, int nb,
                                                           uint32_t* __restrict__ rel, uint32_t* __restrict__ firstRel) {
    uint32_t p = blockIdx.x * kThreads + threadIdx.x;
    if (p >= np) return;
    rel[p] = flag[p] ? 1u : 0u;
    if (!flag[p]) return;
    uint32_t i, j;
    ccdBodies(bodies, pairs[p], nb, i, j);
    if (isDynamic(bodies[i])) atomicMin(&firstRel[i], p);
    if (isDynamic(bodies[j])) atomicMin(&firstRel[j], p);
}

constexpr int kCcdRounds = 8;

// One closure round.  `prevAdded`: pairs the previous round added (null in
// the first round); a round after one that added none has nothing to do.
__global__ __launch_bounds__(kThreads) void kCcdRelRound(const vx_body* __restrict__ bodies,
                                                         const uint64_t* __restrict__ pairs,
                                                         const uint32_t* __restrict__ counts, uint32_t np, int nb,
                                                         uint32_t* firstRel, uint32_t* __restrict__ rel,
                                                         uint32_t* __restrict__ changed,
                                                         const uint32_t* __restrict__ prevAdded) {
    if (prevAdded && *prevAdded == 0) return;
    uint32_t p = blockIdx.x * kThreads + threadIdx.x;
    if (p >= np || rel[p] || !counts[p]) return;
    uint32_t i, j;
    ccdBodies(bodies, pairs[p], nb, i, j);
    const bool di = isDynamic(bodies[i]), dj = isDynamic(bodies[j]);
    const bool behind = (di && __hip_atomic_load(&firstRel[i], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) < p) ||
                        (dj && __hip_atomic_load(&firstRel[j], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) < p);
    if (behind) {
        rel[p] = 1u;
        atomicAdd(changed, 1u);
        if (di) atomicMin(&firstRel[i], p);
        if (dj) atomicMin(&firstRel[j], p);
    }
}

// Fallback when the closure has not settled after kCcdRounds rounds (the
// last one still added pairs): every pair that produced contacts.
__global__ __launch_bounds__(kThreads) void kCcdRelAll(const uint32_t* __restrict__ counts, uint32_t np,
                                                       uint32_t* __restrict__ rel,
                                                       const uint32_t* __restrict__ lastAdded) {
    if (*lastAdded == 0) return;
    uint32_t p = blockIdx.x * kThreads + threadIdx.x;
    if (p < np) rel[p] = counts[p] ? 1u : 0u;
}

// Relevant pairs listed in increasing pair order.
__global__ __launch_bounds__(kThreads) void kCcdRelList(const uint32_t* __restrict__ rel,
                                                        const uint32_t* __restrict__ relPos, uint32_t np,
                                                        uint32_t* __restrict__ list) {
    uint32_t p = blockIdx.x * kThreads + threadIdx.x;
    if (p < np && rel[p]) list[relPos[p]] = p;
}

// One (body << pb | pair) entry per dynamic body of every relevant pair;
// sorted, they give each pair its predecessor on each body.
__global__ __launch_bounds__(kThreads) void kCcdEntries(const vx_body* __restrict__ bodies,
                                                        const uint64_t* __restrict__ pairs,
                                                        const uint32_t* __restrict__ list,
                                                        const uint32_t* __restrict__ nrDev, int nb,
                                                        int pbits, uint64_t* __restrict__ keys,
                                                        uint32_t* __restrict__ vals, uint32_t* __restrict__ counter) {
    const uint32_t nr = *nrDev;
    uint32_t t = blockIdx.x * kThreads + threadIdx.x;
    bool s0 = false, s1 = false;
    uint32_t i = 0, j = 0, p = 0;
    if (t < nr) {
        p = list[t];
        ccdBodies(bodies, pairs[p], nb, i, j);
        s0 = isDynamic(bodies[i]);
        s1 = isDynamic(bodies[j]);
    }
    uint32_t at = waveReserve((s0 ? 1u : 0u) + (s1 ? 1u : 0u), counter);
    if (s0) { keys[at] = ((uint64_t)i << pbits) | p; vals[at] = p << 1; ++at; }
    if (s1) { keys[at] = ((uint64_t)j << pbits) | p; vals[at] = (p << 1) | 1u; }
}

// The net in reference order by dataflow across workgroups, over the
// relevant pairs only: a pair waits for the previous relevant pairs on its
// dynamic bodies (one atomic word each: done bit and a "moved" bit per body),
// re-examines itself only when flagged or when one of its bodies was rewound
// earlier, and passes the moved bits on.  Body data is published with
// release/acquire at device scope only along chains that carry a move; all
// other hand-offs are relaxed.  Consecutive list entries go to different
// wavefronts (entry t to wavefront t mod W), because a rescue is long
// divergent GJK work and the flagged pairs sit close together in the list;
// each lane walks its entries in increasing order and the grid is sized to be
// co-resident, so the lowest unfinished pair is always running and ready; a
// spin budget bounds waits.  The kernel runs one wavefront per SIMD, so one
// workgroup of four fits a CU: the grid takes three CUs in four (more
// wavefronts share fewer entries each: 128 workgroups took 2.27 ms per step
// of debris, 192 on 256 CUs 2.14 ms), keeping a quarter free.
constexpr uint32_t kCcdGridPerCu4 = 3;  // workgroups per four CUs
constexpr int kCoopRescues = 2;  // at most this many ready rescues share the wavefront

__global__ __launch_bounds__(kThreads) void kCcdDataflow(vx_body* bodies, const vx_vec3* __restrict__ prevPos,
                                                         const vx_quat* __restrict__ prevQ, Soup soup, float dt,
                                                         const uint64_t* __restrict__ pairs,
                                                         const uint32_t* __restrict__ list,
                                                         const uint32_t* __restrict__ nrDev, int nb,
                                                         const uint8_t* __restrict__ flag,
                                                         const float4* __restrict__ endGap,
                                                         const uint32_t* __restrict__ pred, uint32_t* word,
                                                         uint32_t* __restrict__ examined,
                                                         uint32_t* __restrict__ rescued, uint32_t* __restrict__ error,
                                                         const uint8_t* __restrict__ toiDone) {
    const uint32_t nr = *nrDev;
    const uint32_t waves = gridDim.x * (kThreads / 64);
    const uint32_t wave = blockIdx.x * (kThreads / 64) + threadIdx.x / 64, lane = threadIdx.x % 64;
    const uint32_t stride = 64 * waves;
    uint32_t t = wave + waves * lane;
    uint32_t spins = 0;
    bool active = t < nr;
    // The wavefront loops while any lane has entries left, so it is whole at
    // the rescues below.
    while (__any(active)) {
        bool ready = false, exam = false, movedA = false, movedB = false;
        uint32_t p = 0, i = 0, j = 0;
        if (active) {
            p = list[t];
            const uint32_t pa = pred[2 * p], pb = pred[2 * p + 1];
            const uint32_t wa = pa == kNone ? 1u : __hip_atomic_load(&word[pa >> 1], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            const uint32_t wb = pb == kNone ? 1u : __hip_atomic_load(&word[pb >> 1], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            ready = (wa & 1u) && (wb & 1u);
            if (ready) {
                movedA = pa != kNone && ((wa >> (1 + (pa & 1))) & 1u);
                movedB = pb != kNone && ((wb >> (1 + (pb & 1))) & 1u);
                if (movedA || movedB) __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
                if (flag[p] || movedA || movedB) {
                    ccdBodies(bodies, pairs[p], nb, i, j);
                    exam = !(toiDone && (toiDone[i] || toiDone[j]));
                }
            } else if (++spins > (1u << 24)) {
                atomicOr(error, 1u);
                active = false;
            }
        }
        // Pairs ready together share no dynamic body (a pair waits for every
        // earlier relevant pair on its bodies), so their rescues commute.  One
        // or two of them (the usual case: consecutive entries sit on different
        // wavefronts) run one after the other on the whole wavefront, which
        // shares each hull support across the lanes; more run one per lane.
        bool r = false;
        uint64_t need = __ballot(exam);
        if (__popcll(need) <= kCoopRescues) {
            while (need) {
                const int owner = __ffsll((unsigned long long)need) - 1;
                const uint32_t po = (uint32_t)__builtin_amdgcn_readlane((int)p, owner);
                const uint32_t io = (uint32_t)__builtin_amdgcn_readlane((int)i, owner);
                const uint32_t jo = (uint32_t)__builtin_amdgcn_readlane((int)j, owner);
                const bool movedO = __builtin_amdgcn_readlane((movedA || movedB) ? 1 : 0, owner) != 0;
                const bool ro = ccdRescue(bodies, prevPos, prevQ, io, jo, soup, dt, movedO ? nullptr : &endGap[po], true,
                                          lane == (uint32_t)owner);
                if (lane == (uint32_t)owner) r = ro;
                need &= need - 1;
            }
        } else if (exam) {
            r = ccdRescue(bodies, prevPos, prevQ, i, j, soup, dt, (movedA || movedB) ? nullptr : &endGap[p]);
        }
        if (ready) {
            if (exam) {
                atomicAdd(examined, 1u);
                if (r) atomicAdd(rescued, 1u);
            }
            const uint32_t out = 1u | ((movedA || r) ? 2u : 0u) | ((movedB || r) ? 4u : 0u);
            if (out != 1u) __hip_atomic_store(&word[p], out, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
            else __hip_atomic_store(&word[p], out, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            t += stride;
            spins = 0;
            active = t < nr;
        }
    }
}

}  // namespace vxk
