// The multi-TOI event scheduler for MotionQuality::High bodies on the device:
// processHighToi (world.cpp:3019-3142) with findToi (world.cpp:2966-3017).
//
// Eligible bodies (High, continuous, fast enough at the step start, and only
// touching static or other eligible bodies) are rewound to the step start and
// replayed on one shared timeline: each event is the earliest time of impact
// over every (eligible, eligible-or-static) pair, ties broken by body indices;
// every eligible body advances to it, the pair exchanges one impulse and a
// small positional separation, and the search repeats on the time left.
//
//   kToiEligible    eligibility at the step start (world.cpp:3028-3038)
//   kToiContactPass the one-sided-rewind exclusion, sequential in contact
//                   order as in the reference (world.cpp:3042-3046)
//   kToiRewind      rewind to the step start, compaction of the eligible and
//                   static bodies
//   kToiPairs       findToi of every candidate pair against the current state,
//                   minimum time by atomicMin on its float bits (times are >= 0)
//   kToiPick        the smallest (a, b) among the pairs at that time
//   kToiAdvance     every eligible body advances by the event time
//   kToiImpulse     the event's impulse and separation (world.cpp:3095-3130)
#pragma once
#include "vx_ccd.hpp"

namespace vxk {

struct ToiSettings {  // WorldMultiToiSettings (ccd.h:111-122)
    uint32_t enabled, splitting, maxPerBody, maxTotal;
    float velocityFloor, bias;
};

// Device-side scheduler state (one per world).
struct ToiState {
    unsigned long long bestPair;  // (a << 32) | b of the event
    float remaining;
    uint32_t bestBits;            // float bits of the event time; 0xffffffff: none
    uint32_t prevA, prevB, total, eligible, statics;
    float normal[3], point[3];    // the event's normal and contact point
};

__global__ __launch_bounds__(kThreads) void kToiEligible(const vx_body* __restrict__ bodies,
                                                         const vx_body* __restrict__ starts, uint32_t n,
                                                         ToiSettings ts, uint8_t* __restrict__ eligible) {
    uint32_t i = blockIdx.x * kThreads + threadIdx.x;
    if (i >= n) return;
    const vx_body& s = starts[i];
    const float floorV = fmx(s.ccdTuning.minVelocityForCCD, ts.velocityFloor);
    eligible[i] = isDynamic(s) && !asleepFully(bodies[i]) && s.ccdTuning.quality == VX_QUALITY_HIGH &&
                  s.ccdTuning.enableContinuous && maxPointSpeed(s) >= floorV;
}

// A medium-quality moving counterpart keeps a High body on the regular path;
// one lane applies the rule in contact order to the contacts that touch an
// eligible body (eligibility only ever drops, so the filter is a superset).
__global__ __launch_bounds__(kWakeThreads) void kToiContactPass(const vx_body* __restrict__ bodies,
                                                                const vx_contact* __restrict__ cs, uint32_t m,
                                                                uint8_t* eligible) {
    __shared__ uint8_t relevant[kWakeThreads];
    for (uint32_t base = 0; base < m; base += kWakeThreads) {
        const uint32_t k = base + threadIdx.x;
        bool r = false;
        if (k < m) r = eligible[cs[k].a] || eligible[cs[k].b];
        relevant[threadIdx.x] = r ? 1 : 0;
        if (__syncthreads_or(r) && threadIdx.x == 0) {
            const uint32_t end = m - base < (uint32_t)kWakeThreads ? m - base : (uint32_t)kWakeThreads;
            for (uint32_t j = 0; j < end; ++j) {
                if (!relevant[j]) continue;
                const uint32_t a = cs[base + j].a, b = cs[base + j].b;
                if (eligible[a] && isDynamic(bodies[b]) && !eligible[b]) eligible[a] = 0;
                if (eligible[b] && isDynamic(bodies[a]) && !eligible[a]) eligible[b] = 0;
            }
        }
        __syncthreads();
    }
}

__global__ __launch_bounds__(kThreads) void kToiRewind(vx_body* __restrict__ bodies, const vx_body* __restrict__ starts,
                                                       uint32_t n, const uint8_t* __restrict__ eligible,
                                                       uint32_t* __restrict__ elist, uint32_t* __restrict__ slist,
                                                       ToiState* st) {
    uint32_t i = blockIdx.x * kThreads + threadIdx.x;
    if (i >= n) return;
    if (eligible[i]) {
        const vx_body& s = starts[i];
        bodies[i].position = s.position;
        bodies[i].orientation = s.orientation;
        bodies[i].velocity = s.velocity;
        bodies[i].angularVelocity = s.angularVelocity;
        elist[atomicAdd(&st->eligible, 1u)] = i;
    } else if (isStatic(bodies[i])) {
        slist[atomicAdd(&st->statics, 1u)] = i;
    }
}

// findToi (world.cpp:2966-3017) against the bodies' current state.
__device__ inline bool findToi(const vx_body& a0, const vx_body& b0, float dt, float bias, const Soup& soup,
                               float& toi, V3& normal, V3& point) {
    const float speedBound = maxPointSpeed(a0) + maxPointSpeed(b0);
    if (speedBound <= 1e-8f) return false;
    const float search = speedBound * dt + a0.radius + length(v3(a0.halfExtents)) + a0.capsuleHalfHeight + 0.1f;
    float lower = 0.0f, upper = 0.0f;
    Gap sample = gapAt(a0, v3(a0.position), q4(a0.orientation), b0, v3(b0.position), q4(b0.orientation), soup, search);
    if (sample.gap <= bias) {
        toi = 0.0f;
        normal = sample.normal;
    } else {
        bool hit = false;
        for (int it = 0; it < 24 && lower < dt; ++it) {
            const float increment = fmx(1e-6f, (sample.gap - bias) / speedBound);
            upper = fmn(dt, lower + increment);
            vx_body a = a0, b = b0;
            advanceTransform(a, upper);
            advanceTransform(b, upper);
            sample = gapAt(a, v3(a.position), q4(a.orientation), b, v3(b.position), q4(b.orientation), soup, search);
            if (sample.gap <= bias) { hit = true; break; }
            lower = upper;
            if (upper >= dt) break;
        }
        if (!hit) return false;
        for (int it = 0; it < 16; ++it) {
            const float middle = (lower + upper) * 0.5f;
            vx_body a = a0, b = b0;
            advanceTransform(a, middle);
            advanceTransform(b, middle);
            const Gap mg = gapAt(a, v3(a.position), q4(a.orientation), b, v3(b.position), q4(b.orientation), soup,
                                 search);
            if (mg.gap <= bias) { upper = middle; sample = mg; }
            else lower = middle;
        }
        toi = upper;
        normal = sample.normal;
    }
    vx_body a = a0;
    advanceTransform(a, toi);
    const Convex cv = makeConvex(a, soup);
    point = cv.support(-normal) - normal * cv.radius;
    return true;
}

// World::queryMultiToi (world.cpp:3146-3185) for one source body against the
// resident state: one lane per other body.  An ineligible source (not
// dynamic, locked, continuous off, Low quality, or slower than the larger of
// its own CCD floor and the settings' floor) has no hits.  A hit leaves its
// time of impact bits in toiBits (kNone otherwise) and (normal, fraction),
// (point, toi) in hit[2 o], hit[2 o + 1].
__global__ __launch_bounds__(kThreads) void kToiQuery(const vx_body* __restrict__ bodies, uint32_t n, Soup soup,
                                                      uint32_t src, float dt, ToiSettings ts,
                                                      uint32_t* __restrict__ toiBits, float4* __restrict__ hit) {
    const uint32_t o = blockIdx.x * kThreads + threadIdx.x;
    if (o >= n) return;
    const vx_body& s = bodies[src];
    const float floorSpeed = s.ccdTuning.minVelocityForCCD < ts.velocityFloor ? ts.velocityFloor
                                                                               : s.ccdTuning.minVelocityForCCD;
    const bool eligible = isDynamic(s) && !isLocked(s) && s.ccdTuning.enableContinuous &&
                          s.ccdTuning.quality != VX_QUALITY_LOW && !(maxPointSpeed(s) < floorSpeed);
    uint32_t bits = kNone;
    if (eligible && o != src) {
        const vx_body& b = bodies[o];
        if (canCollide(s, b) && !(isStatic(b) && isStatic(s))) {
            float toi;
            V3 nrm, pt;
            if (findToi(s, b, dt, ts.bias, soup, toi, nrm, pt)) {
                bits = __float_as_uint(toi);
                hit[2 * o] = make_float4(nrm.x, nrm.y, nrm.z, toi / dt);
                hit[2 * o + 1] = make_float4(pt.x, pt.y, pt.z, toi);
            }
        }
    }
    toiBits[o] = bits;
}

__device__ inline void toiPair(uint32_t p, uint32_t ne, uint32_t width, const uint32_t* elist, const uint32_t* slist,
                               uint32_t& a, uint32_t& b, bool& dynamicPair) {
    const uint32_t ai = p / width, bj = p % width;
    a = elist[ai];
    dynamicPair = bj < ne;
    b = dynamicPair ? elist[bj] : slist[bj - ne];
}

// Candidates (world.cpp:3072-3082): a eligible with events left; b eligible
// above a (with events left) or static; b no sensor; filters pass.
__global__ __launch_bounds__(kThreads) void kToiPairs(const vx_body* __restrict__ bodies, Soup soup, ToiSettings ts,
                                                      const uint32_t* __restrict__ elist,
                                                      const uint32_t* __restrict__ slist,
                                                      const uint32_t* __restrict__ bodyEvents, uint32_t ne,
                                                      uint32_t ns, ToiState* st, uint32_t* __restrict__ toiBits,
                                                      float4* __restrict__ hit) {
    const uint32_t width = ne + ns, p = blockIdx.x * kThreads + threadIdx.x;
    if (p >= ne * width) return;
    uint32_t a, b;
    bool dyn;
    toiPair(p, ne, width, elist, slist, a, b, dyn);
    uint32_t bits = 0xffffffffu;
    const float remaining = st->remaining;
    if (a != b && bodyEvents[a] < ts.maxPerBody && !(dyn && b < a) && !isSensor(bodies[b]) &&
        canCollide(bodies[a], bodies[b]) && !(dyn && bodyEvents[b] >= ts.maxPerBody)) {
        float toi;
        V3 nrm, pt;
        if (findToi(bodies[a], bodies[b], remaining, ts.bias, soup, toi, nrm, pt) &&
            !(a == st->prevA && b == st->prevB && toi <= 1e-6f)) {
            bits = __float_as_uint(toi);
            atomicMin(&st->bestBits, bits);
            hit[2 * p] = make_float4(nrm.x, nrm.y, nrm.z, 0.0f);
            hit[2 * p + 1] = make_float4(pt.x, pt.y, pt.z, 0.0f);
        }
    }
    toiBits[p] = bits;
}

__global__ __launch_bounds__(kThreads) void kToiPick(const uint32_t* __restrict__ elist,
                                                     const uint32_t* __restrict__ slist, uint32_t ne, uint32_t ns,
                                                     const uint32_t* __restrict__ toiBits, ToiState* st) {
    const uint32_t width = ne + ns, p = blockIdx.x * kThreads + threadIdx.x;
    if (p >= ne * width || toiBits[p] != st->bestBits || toiBits[p] == 0xffffffffu) return;
    uint32_t a, b;
    bool dyn;
    toiPair(p, ne, width, elist, slist, a, b, dyn);
    atomicMin(&st->bestPair, ((unsigned long long)a << 32) | b);
}

// The pair at the event time's normal and point.
__global__ __launch_bounds__(kThreads) void kToiSelect(const uint32_t* __restrict__ elist,
                                                       const uint32_t* __restrict__ slist, uint32_t ne, uint32_t ns,
                                                       const uint32_t* __restrict__ toiBits,
                                                       const float4* __restrict__ hit, ToiState* st) {
    const uint32_t width = ne + ns, p = blockIdx.x * kThreads + threadIdx.x;
    if (p >= ne * width || toiBits[p] != st->bestBits || toiBits[p] == 0xffffffffu) return;
    uint32_t a, b;
    bool dyn;
    toiPair(p, ne, width, elist, slist, a, b, dyn);
    if ((((unsigned long long)a << 32) | b) != st->bestPair) return;
    const float4 nv = hit[2 * p], pv = hit[2 * p + 1];
    st->normal[0] = nv.x; st->normal[1] = nv.y; st->normal[2] = nv.z;
    st->point[0] = pv.x; st->point[1] = pv.y; st->point[2] = pv.z;
}

// Every eligible body advances by the event time (or, after the last event,
// by the time left).
__global__ __launch_bounds__(kThreads) void kToiAdvance(vx_body* __restrict__ bodies,
                                                        const uint32_t* __restrict__ elist, uint32_t ne,
                                                        const ToiState* st, int finish) {
    const uint32_t i = blockIdx.x * kThreads + threadIdx.x;
    if (i >= ne) return;
    if (!finish && st->bestBits == 0xffffffffu) return;
    const float h = finish ? st->remaining : __uint_as_float(st->bestBits);
    vx_body b = bodies[elist[i]];
    advanceTransform(b, h);
    bodies[elist[i]] = b;
}

// The event itself (world.cpp:3092-3130), then the scheduler's bookkeeping.
__global__ void kToiImpulse(vx_body* bodies, ToiSettings ts, const uint8_t* __restrict__ eligible,
                            uint32_t* bodyEvents, ToiState* st) {
    if (threadIdx.x != 0 || blockIdx.x != 0 || st->bestBits == 0xffffffffu) return;
    const float toi = __uint_as_float(st->bestBits);
    const uint32_t ia = (uint32_t)(st->bestPair >> 32), ib = (uint32_t)st->bestPair;
    vx_body& a = bodies[ia];
    vx_body& b = bodies[ib];
    const V3 n{st->normal[0], st->normal[1], st->normal[2]}, point{st->point[0], st->point[1], st->point[2]};
    const V3 ra = point - v3(a.position), rb = point - v3(b.position);
    const float vn = dot(v3(a.velocity) + cross(v3(a.angularVelocity), ra) - v3(b.velocity) -
                             cross(v3(b.angularVelocity), rb),
                         n);
    const float angularMassA = dot(cross(invInertiaMul(a, cross(ra, n)), ra), n);
    const float angularMassB = dot(cross(invInertiaMul(b, cross(rb, n)), rb), n);
    const float mass = solverInvMass(a) + solverInvMass(b) + angularMassA + angularMassB;
    if (vn < 0.0f && mass > 1e-8f) {
        const float rest = vn < -1.0f ? combine(a.restitution, b.restitution, a.restitutionCombine, b.restitutionCombine)
                                      : 0.0f;
        const V3 impulse = n * (-(1.0f + rest) * vn / mass);
        a.velocity = raw(v3(a.velocity) + impulse * solverInvMass(a));
        a.angularVelocity = raw(v3(a.angularVelocity) + invInertiaMul(a, cross(ra, impulse)));
        b.velocity = raw(v3(b.velocity) - impulse * solverInvMass(b));
        b.angularVelocity = raw(v3(b.angularVelocity) - invInertiaMul(b, cross(rb, impulse)));
    }
    const float correction = fmx(2.0f * ts.bias, 1e-4f);
    const float inverseMass = solverInvMass(a) + solverInvMass(b);
    if (inverseMass > 0.0f) {
        a.position = raw(v3(a.position) + n * (correction * solverInvMass(a) / inverseMass));
        b.position = raw(v3(b.position) - n * (correction * solverInvMass(b) / inverseMass));
    }
    st->remaining -= toi;
    st->prevA = ia;
    st->prevB = ib;
    ++bodyEvents[ia];
    if (eligible[ib]) ++bodyEvents[ib];
    ++st->total;
}

}  // namespace vxk
