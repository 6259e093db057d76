// Island sleeping on the device: SleepManager::update (sleep.cpp:50-325) as
// World::updateSleeping runs it after the contact events (world.cpp:4492).
//
//   kSleepInit         union-find roots, per-body contact flags, island timers
//   kSleepContacts     islands over dynamic-dynamic non-sensor contacts
//                      (sleep.cpp:85-91) and contact stability flags
//                      (sleep.cpp:107-127)
//   kSleepBodies       stability counters (sleep.cpp:129-138), calm test and
//                      sleep timers (sleep.cpp:152-182), island minimum timer
//                      (sleep.cpp:191-210)
//   kSleepTransitions  island-wide Asleep / Drowsy transitions
//                      (sleep.cpp:221-252) and the per-body state tallies
//   kSleepIslands      island tallies by state (sleep.cpp:264-318)
//
// The island partition does not depend on the order in which unions happen,
// nor do the minimum timers; every per-body quantity is computed with the
// reference's float expressions in the same order.  Islands join over
// contacts and joints (kSleepJoints, sleep.cpp:92-95).
#pragma once
#include "vx_solve.hpp"

namespace vxk {

struct SleepParams {
    float linTolSq, angTolSq, accTolSq;  // squared thresholds (sleep.cpp:149-151)
    float stabilityThreshold;
    float timeToSleep, timeToDrowsy;     // timeToDrowsy already resolved for gradual sleep off
    float dt;
    uint32_t stabilityFrames;            // uint16_t(contactStabilityFrames)
    uint32_t gradual, stability;
};

// Per-body contact flags of the stability pass.
enum : uint32_t { kSlHasContact = 1u, kSlUnstable = 2u };
// Island state bits (sleep.cpp:283-289).
enum : uint32_t { kSlRootAwake = 1u, kSlRootDrowsy = 2u };
constexpr uint32_t kFloat1e30Bits = 0x7149f2cau;  // 1e30f, the island timer start (sleep.cpp:191)

__device__ inline bool asleepDrowsy(const vx_body& b) { return b.asleep == 2; }  // sleep.h:70

__global__ __launch_bounds__(kThreads) void kSleepInit(uint32_t n, uint32_t* __restrict__ parent,
                                                       uint32_t* __restrict__ flags, uint32_t* __restrict__ islandMin,
                                                       uint32_t* __restrict__ rootState) {
    uint32_t i = blockIdx.x * kThreads + threadIdx.x;
    if (i >= n) return;
    parent[i] = i;
    flags[i] = 0;
    islandMin[i] = kFloat1e30Bits;
    rootState[i] = 0;
}

__device__ inline void ufUnite(uint32_t* parent, uint32_t a, uint32_t b) {
    for (;;) {
        a = ufFind(parent, a);
        b = ufFind(parent, b);
        if (a == b) return;
        if (a < b) { uint32_t t = a; a = b; b = t; }
        uint32_t expected = a;
        if (__hip_atomic_compare_exchange_strong(&parent[a], &expected, b, __ATOMIC_RELAXED, __ATOMIC_RELAXED,
                                                 __HIP_MEMORY_SCOPE_AGENT))
            return;
    }
}

__global__ __launch_bounds__(kThreads) void kSleepContacts(const vx_body* __restrict__ bodies, uint32_t n,
                                                           const vx_contact* __restrict__ cs, uint32_t m,
                                                           SleepParams sp, uint32_t* parent,
                                                           uint32_t* __restrict__ flags) {
    uint32_t k = blockIdx.x * kThreads + threadIdx.x;
    if (k >= m) return;
    const vx_contact& c = cs[k];
    if (c.a >= n || c.b >= n) return;
    const vx_body& a = bodies[c.a];
    const vx_body& b = bodies[c.b];
    if (isSensor(a) || isSensor(b)) return;
    const bool dynA = isDynamic(a), dynB = isDynamic(b);
    if (dynA && dynB) ufUnite(parent, c.a, c.b);
    if (!sp.stability) return;
    const V3 p = v3(c.point);
    const V3 va = v3(a.velocity) + cross(v3(a.angularVelocity), p - v3(a.position));
    const V3 vb = v3(b.velocity) + cross(v3(b.angularVelocity), p - v3(b.position));
    const float rel = length(va - vb);
    const uint32_t f = kSlHasContact | (rel > sp.stabilityThreshold ? kSlUnstable : 0u);
    if (dynA) atomicOr(&flags[c.a], f);
    if (dynB) atomicOr(&flags[c.b], f);
}

// Joints join islands too (sleep.cpp:92-95).
__global__ __launch_bounds__(kThreads) void kSleepJoints(const vx_body* __restrict__ bodies, uint32_t n,
                                                         const vx_joint* __restrict__ joints, uint32_t nj,
                                                         uint32_t* parent) {
    uint32_t k = blockIdx.x * kThreads + threadIdx.x;
    if (k >= nj) return;
    const uint32_t a = joints[k].a, b = joints[k].b;
    if (a >= n || b >= n) return;
    if (isDynamic(bodies[a]) && isDynamic(bodies[b])) ufUnite(parent, a, b);
}

__global__ __launch_bounds__(kThreads) void kSleepBodies(vx_body* __restrict__ bodies, uint32_t n, SleepParams sp,
                                                         uint32_t* parent, const uint32_t* __restrict__ flags,
                                                         uint16_t* __restrict__ stability,
                                                         uint32_t* __restrict__ islandMin, uint32_t* __restrict__ wakeCount) {
    uint32_t i = blockIdx.x * kThreads + threadIdx.x;
    bool woke = false;
    if (i < n && isDynamic(bodies[i])) {
        vx_body& b = bodies[i];
        if (sp.stability) {
            const uint32_t f = flags[i];
            const uint16_t st = stability[i];
            stability[i] = (f & kSlHasContact) && !(f & kSlUnstable) ? (uint16_t)(st < 0xffffu ? st + 1 : st) : 0;
        }
        if (!asleepFully(b)) {
            if (!b.enableSleep) {
                // A body that never sleeps keeps its whole island awake.
                atomicMin(&islandMin[ufFind(parent, i)], 0u);
            } else {
                const float linSq = lengthSq(v3(b.velocity));
                const float angSq = lengthSq(v3(b.angularVelocity));
                const float accSq = b.invMass > 0.0f ? lengthSq(v3(b.force)) * b.invMass * b.invMass : 0.0f;
                bool calm = linSq < sp.linTolSq && angSq < sp.angTolSq && accSq < sp.accTolSq;
                if (calm && sp.stability) calm = stability[i] >= (uint16_t)sp.stabilityFrames;
                if (calm) {
                    b.sleepTimer += sp.dt;
                } else {
                    b.sleepTimer = 0.0f;
                    if (asleepDrowsy(b)) { b.asleep = 0; woke = true; }
                }
                atomicMin(&islandMin[ufFind(parent, i)], __float_as_uint(b.sleepTimer));
            }
        }
    }
    waveReserve(woke ? 1u : 0u, wakeCount);
}

__global__ __launch_bounds__(kThreads) void kSleepTransitions(vx_body* __restrict__ bodies, uint32_t n, SleepParams sp,
                                                              uint32_t* parent, const uint32_t* __restrict__ islandMin,
                                                              uint32_t* __restrict__ rootState,
                                                              uint32_t* __restrict__ counters /* 6 */) {
    uint32_t i = blockIdx.x * kThreads + threadIdx.x;
    bool dyn = false, toSleep = false, toDrowsy = false;
    uint8_t state = 0;
    if (i < n && isDynamic(bodies[i])) {
        dyn = true;
        vx_body& b = bodies[i];
        const uint32_t root = ufFind(parent, i);
        if (!asleepFully(b) && b.enableSleep) {
            const float minTimer = __uint_as_float(islandMin[root]);
            if (minTimer >= sp.timeToSleep) {
                b.asleep = 1;
                b.velocity = {0, 0, 0};
                b.angularVelocity = {0, 0, 0};
                b.sleepTimer = 0.0f;
                toSleep = true;
            } else if (sp.gradual && minTimer >= sp.timeToDrowsy && !asleepDrowsy(b)) {
                b.asleep = 2;
                toDrowsy = true;
            }
        }
        state = b.asleep;
        if (state == 0) atomicOr(&rootState[root], kSlRootAwake);
        else if (state != 1) atomicOr(&rootState[root], kSlRootDrowsy);
    }
    waveReserve(dyn ? 1u : 0u, counters + 0);
    waveReserve(dyn && state == 0 ? 1u : 0u, counters + 1);
    waveReserve(dyn && state != 0 && state != 1 ? 1u : 0u, counters + 2);
    waveReserve(dyn && state == 1 ? 1u : 0u, counters + 3);
    waveReserve(toSleep ? 1u : 0u, counters + 4);
    waveReserve(toDrowsy ? 1u : 0u, counters + 5);
}

// Islands are the roots of dynamic bodies (a dynamic body's root is dynamic).
__global__ __launch_bounds__(kThreads) void kSleepIslands(const vx_body* __restrict__ bodies, uint32_t n,
                                                          const uint32_t* __restrict__ parent,
                                                          const uint32_t* __restrict__ rootState,
                                                          uint32_t* __restrict__ counters /* 4 */) {
    uint32_t i = blockIdx.x * kThreads + threadIdx.x;
    const bool root = i < n && isDynamic(bodies[i]) && parent[i] == i;
    const uint32_t st = root ? rootState[i] : 0u;
    waveReserve(root ? 1u : 0u, counters + 0);
    waveReserve(root && !st ? 1u : 0u, counters + 1);                                   // asleep
    waveReserve(root && !(st & kSlRootAwake) && (st & kSlRootDrowsy) ? 1u : 0u, counters + 2);  // drowsy
    waveReserve(root && (st & kSlRootAwake) ? 1u : 0u, counters + 3);                   // awake
}

}  // namespace vxk
