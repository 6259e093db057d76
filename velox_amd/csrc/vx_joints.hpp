// Joints on the device, as the reference's CUDA backend runs them inside its
// device substeps (backend_cuda.cu:691-761, 800-891) and as World::stepImpl
// treats them around it (world.cpp:4045-4066, 3765-3783):
//
//   kJointPairsOff    contacts of joint-connected pairs (collideConnected off)
//                     are dropped, order kept, and their pairs leave the CCD net
//   kJointWake        a joint between a sleeping and a waking body wakes both,
//                     in joint order (one lane; only when a body sleeps)
//   kJointReset       accumulated impulses cleared each substep (broken joints
//                     kept after the first substep)
//   kJointPacks       packs of whole joint-graph components, one workgroup
//                     each: 8 passes over the pack's colours, a barrier
//                     between colours
//   kJointColor       the joints of one colour, one lane each (colours never
//                     share a dynamic body), 8 passes over the colours
//   kJointMarkBroken  reaction force / torque over the substep against the
//                     joint's thresholds
//   kJointBroken      World::finishBrokenJoints' break test, all joints at once
//   kJointFinish      World::finishBrokenJoints: broken joints removed from the
//                     end, each swapped with the last, with a break event
#pragma once
#include "vxd_joint.hpp"
#include "vx_kernels.hpp"

namespace vxk {

// One joint break (World::jointBreakEvents_, world.cpp:3776-3780): the joint's
// dense index when it was removed, its bodies, reaction force and torque.
struct JointBreak { uint32_t index, a, b; float force, torque; };

__global__ __launch_bounds__(kThreads) void kJointContactsOff(const vx_contact* __restrict__ cs, uint32_t m,
                                                              const uint64_t* __restrict__ excluded, uint32_t ne,
                                                              uint32_t* __restrict__ keep) {
    uint32_t k = blockIdx.x * kThreads + threadIdx.x;
    if (k >= m) return;
    const uint32_t lo = cs[k].a < cs[k].b ? cs[k].a : cs[k].b, hi = cs[k].a < cs[k].b ? cs[k].b : cs[k].a;
    const uint64_t key = (uint64_t)lo << 32 | hi;
    const uint32_t at = lowerBound64(excluded, ne, key);
    keep[k] = at < ne && excluded[at] == key ? 0u : 1u;
}

__global__ __launch_bounds__(kThreads) void kJointContactsCompact(const vx_contact* __restrict__ in, uint32_t m,
                                                                  const uint32_t* __restrict__ keep,
                                                                  const uint32_t* __restrict__ pos,
                                                                  vx_contact* __restrict__ out) {
    uint32_t k = blockIdx.x * kThreads + threadIdx.x;
    if (k < m && keep[k]) out[pos[k]] = in[k];
}

// The CCD net's pair list is the contacts' pairs; an excluded pair leaves it.
__global__ __launch_bounds__(kThreads) void kJointPairsOff(const uint64_t* __restrict__ pairs, uint32_t np, int nb,
                                                           const uint64_t* __restrict__ excluded, uint32_t ne,
                                                           uint32_t* __restrict__ counts) {
    uint32_t p = blockIdx.x * kThreads + threadIdx.x;
    if (p >= np || !counts[p]) return;
    const uint32_t i = (uint32_t)(pairs[p] >> nb), j = (uint32_t)(pairs[p] & ((1ull << nb) - 1));
    const uint32_t lo = i < j ? i : j, hi = i < j ? j : i;
    const uint64_t key = (uint64_t)lo << 32 | hi;
    const uint32_t at = lowerBound64(excluded, ne, key);
    if (at < ne && excluded[at] == key) counts[p] = 0;
}

__global__ void kJointWake(vx_body* bodies, const vx_joint* joints, uint32_t nj) {
    if (threadIdx.x || blockIdx.x) return;
    for (uint32_t k = 0; k < nj; ++k) {
        const uint32_t a = joints[k].a, b = joints[k].b;
        if (bodies[a].asleep == bodies[b].asleep) continue;
        bodies[a].asleep = bodies[b].asleep = 0;
        bodies[a].sleepTimer = bodies[b].sleepTimer = 0.0f;
    }
}

__global__ __launch_bounds__(kThreads) void kJointReset(vx_joint* joints, uint32_t nj, int keepBroken) {
    uint32_t k = blockIdx.x * kThreads + threadIdx.x;
    if (k >= nj || (keepBroken && joints[k].broken)) return;
    jointReset(joints[k]);
}

__device__ __attribute__((always_inline)) inline void solveJointAt(vx_body* bodies, vx_joint* joints, uint32_t idx,
                                                                   float h) {
    vx_joint& j = joints[idx];
    if (j.broken) return;
    vx_body a = bodies[j.a], b = bodies[j.b];
    if (asleepFully(a) && asleepFully(b)) return;
    vx_joint jj = j;
    solveJoint(jj, a, b, h);
    j = jj;
    bodies[j.a].velocity = a.velocity; bodies[j.a].angularVelocity = a.angularVelocity;
    bodies[j.b].velocity = b.velocity; bodies[j.b].angularVelocity = b.angularVelocity;
}

__global__ __launch_bounds__(kThreads) void kJointColor(vx_body* bodies, vx_joint* joints,
                                                        const uint32_t* __restrict__ order, uint32_t first,
                                                        uint32_t count, float h) {
    uint32_t k = blockIdx.x * kThreads + threadIdx.x;
    if (k >= count) return;
    solveJointAt(bodies, joints, order[first + k], h);
}

// Joint packs: one workgroup per pack runs all kJointIterations passes over
// the pack's colours, a barrier between colours.  A pack is a union of whole
// components of the joint graph, so no other workgroup writes its dynamic
// bodies, and its joints of one colour share none; each joint therefore sees
// the state the per-colour launches give it.  Static and kinematic endpoints
// are read by several packs and written back unchanged.
constexpr uint32_t kJointPack = 256;        // a pack is closed at this many joints
constexpr uint32_t kJointPackMax = 4096;    // larger packs take the per-colour launches (default)
constexpr int kJointPackThreads = 256;

__global__ __launch_bounds__(kJointPackThreads) void kJointPacks(vx_body* bodies, vx_joint* joints,
                                                                 const uint32_t* __restrict__ order,
                                                                 const uint32_t* __restrict__ runStart,
                                                                 const uint32_t* __restrict__ packFirstRun, float h) {
    const uint32_t r0 = packFirstRun[blockIdx.x], r1 = packFirstRun[blockIdx.x + 1];
    for (int it = 0; it < vxd::kJointIterations; ++it)
        for (uint32_t r = r0; r < r1; ++r) {
            const uint32_t b = runStart[r], e = runStart[r + 1];
            for (uint32_t k = b + threadIdx.x; k < e; k += kJointPackThreads) solveJointAt(bodies, joints, order[k], h);
            __syncthreads();
        }
}

__global__ __launch_bounds__(kThreads) void kJointMarkBroken(vx_joint* joints, uint32_t nj, float h) {
    uint32_t k = blockIdx.x * kThreads + threadIdx.x;
    if (k >= nj || joints[k].broken) return;
    vx_joint& j = joints[k];
    const float force = length(v3(j.reactionLinearImpulse)) / h;
    const float torque = length(v3(j.reactionAngularImpulse)) / h;
    j.broken = force > j.breakForce || torque > j.breakTorque;
}

// The step-end break test of every joint (World::finishBrokenJoints' first
// pass), in parallel; counts the broken joints.
__global__ __launch_bounds__(kThreads) void kJointBroken(vx_joint* joints, uint32_t nj, float h,
                                                         uint32_t* __restrict__ brokenCount) {
    const uint32_t k = blockIdx.x * kThreads + threadIdx.x;
    bool broken = false;
    if (k < nj) {
        vx_joint& j = joints[k];
        const float force = length(v3(j.reactionLinearImpulse)) / h;
        const float torque = length(v3(j.reactionAngularImpulse)) / h;
        broken = j.broken || force > j.breakForce || torque > j.breakTorque;
        j.broken = broken;
    }
    const uint64_t any = __ballot(broken);
    if (any && (threadIdx.x & 63) == 0) atomicAdd(brokenCount, (uint32_t)__popcll(any));
}

// The removal pass, in the reference's order (from the end, each broken
// joint swapped with the last); nothing to walk when none broke.
__global__ void kJointFinish(vx_joint* joints, uint32_t nj, float h, JointBreak* breaks, uint32_t* remaining,
                             uint32_t* breakCount, const uint32_t* __restrict__ brokenCount) {
    if (threadIdx.x || blockIdx.x) return;
    if (*brokenCount == 0) {
        *remaining = nj;
        *breakCount = 0;
        return;
    }
    uint32_t count = nj, nb = 0;
    for (uint32_t i = nj; i-- > 0;) {
        if (!joints[i].broken) continue;
        const vx_joint& j = joints[i];
        breaks[nb++] = JointBreak{i, j.a, j.b, length(v3(j.reactionLinearImpulse)) / h,
                                  length(v3(j.reactionAngularImpulse)) / h};
        if (i != count - 1) joints[i] = joints[count - 1];
        --count;
    }
    *remaining = count;
    *breakCount = nb;
}

}  // namespace vxk
