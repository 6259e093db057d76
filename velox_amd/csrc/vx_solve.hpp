// Colored TGS sweeps on compact per-body arrays (narrowphase.h:130-366).
//
// Within one substep the body poses are fixed (only velocities change between
// sweeps), so everything solveContact derives from poses -- anchors, arms,
// tangents, effective masses, live gap, the inertia frame -- is computed once
// per substep (kPrepareRows) with exactly the operations the reference
// performs on every sweep; the sweep kernel then runs only the
// velocity-dependent rows.  Results are bitwise those of the per-sweep
// recomputation.
#pragma once
#include "vx_kernels.hpp"

namespace vxk {

enum RowFlag : uint32_t { kDynA = 1, kDynB = 2, kRotA = 4, kRotB = 8, kSkip = 16, kSensorRow = 32 };

struct SolveArrays {
    float4* vel;      // per body: linear velocity
    float4* ang;      // per body: angular velocity
    float4* invI;     // per body: inverse principal inertia (xyz), solver inverse mass (w)
    float4* frame;    // per body: principal-inertia frame at the current orientation
    // Per-contact arrays are indexed by the contact's solve position: batched
    // contacts in batch order, then the big-island contacts in colour order.
    float4* rows;     // per contact: 8 float4 (see kPrepareRows)
    uint4* meta;      // per contact: a, b, flags
    float4* imp;      // per contact: 2 float4 (normal, t1, t2, roll1 | roll2, spin)
    float4* side;     // per contact: 2 x (frame, inverse inertia, linear, angular velocity) of A then B
    const uint32_t* pos;  // contact (colour order) -> solve position
};

__device__ inline V3 frameMul(float4 frame, float4 invI, V3 v, bool active) {  // body.h:220-228
    if (!active) return {};
    Q4 f{frame.x, frame.y, frame.z, frame.w};
    V3 local = rotateInv(f, v);
    return rotate(f, V3{local.x * invI.x, local.y * invI.y, local.z * invI.z});
}

// Per substep: compact body arrays from the records.
__global__ __launch_bounds__(kThreads) void kGatherSolveBodies(const vx_body* __restrict__ bodies, uint32_t n,
                                                               SolveArrays sa) {
    uint32_t i = blockIdx.x * kThreads + threadIdx.x;
    if (i >= n) return;
    const vx_body b = bodies[i];
    sa.vel[i] = make_float4(b.velocity.x, b.velocity.y, b.velocity.z, 0.0f);
    sa.ang[i] = make_float4(b.angularVelocity.x, b.angularVelocity.y, b.angularVelocity.z, 0.0f);
    sa.invI[i] = make_float4(b.invInertia.x, b.invInertia.y, b.invInertia.z, solverInvMass(b));
    Q4 f = inertiaFrameAt(b, q4(b.orientation));
    sa.frame[i] = make_float4(f.x, f.y, f.z, f.w);
}

__global__ __launch_bounds__(kThreads) void kScatterVelocities(vx_body* __restrict__ bodies, uint32_t n, SolveArrays sa) {
    uint32_t i = blockIdx.x * kThreads + threadIdx.x;
    if (i >= n) return;
    float4 v = sa.vel[i], w = sa.ang[i];
    bodies[i].velocity = {v.x, v.y, v.z};
    bodies[i].angularVelocity = {w.x, w.y, w.z};
}

__global__ __launch_bounds__(kThreads) void kLoadImpulses(const vx_contact* __restrict__ cs, uint32_t m, float4* imp,
                                                          const uint32_t* __restrict__ pos) {
    uint32_t k = blockIdx.x * kThreads + threadIdx.x;
    if (k >= m) return;
    const vx_contact& c = cs[k];
    const uint32_t p = pos[k];
    imp[2 * p] = make_float4(c.normalImpulse, c.tangentImpulse1, c.tangentImpulse2, c.rollingImpulse1);
    imp[2 * p + 1] = make_float4(c.rollingImpulse2, c.spinningImpulse, 0.0f, 0.0f);
}

__global__ __launch_bounds__(kThreads) void kStoreImpulses(vx_contact* __restrict__ cs, uint32_t m, const float4* imp,
                                                           const uint32_t* __restrict__ pos) {
    uint32_t k = blockIdx.x * kThreads + threadIdx.x;
    if (k >= m) return;
    const uint32_t p = pos[k];
    float4 a = imp[2 * p], b = imp[2 * p + 1];
    cs[k].normalImpulse = a.x; cs[k].tangentImpulse1 = a.y; cs[k].tangentImpulse2 = a.z; cs[k].rollingImpulse1 = a.w;
    cs[k].rollingImpulse2 = b.x; cs[k].spinningImpulse = b.y;
}

// The pose-dependent prefix of solveContact (narrowphase.h:137-224), once per
// substep for every contact (all colours at once: poses are read-only here).
// Row layout:
//   r0 = ra, kN        r1 = rb, k1          r2 = n, k2         r3 = t1, targetBase
//   r4 = t2, liveGap   r5 = vn0, e, mu1, mu2                   r6 = roll, spin, kr1, kr2
//   r7 = ks, kN1, kN2, k12
__device__ inline uint32_t rowFlags(const vx_body& a, const vx_body& b) {
    uint32_t flags = 0;
    if (isDynamic(a)) flags |= kDynA;
    if (isDynamic(b)) flags |= kDynB;
    if (isDynamic(a) && !rotationLocked(a)) flags |= kRotA;
    if (isDynamic(b) && !rotationLocked(b)) flags |= kRotB;
    if (isSensor(a) || isSensor(b)) flags |= kSkip | kSensorRow;
    return flags;
}

// Endpoints and body flags of every contact in colour order (the island
// pass runs on these before solve positions exist).
__global__ __launch_bounds__(kThreads) void kPrepareMeta(const vx_body* __restrict__ bodies,
                                                         const vx_contact* __restrict__ cs, uint32_t m,
                                                         uint4* __restrict__ meta) {
    uint32_t k = blockIdx.x * kThreads + threadIdx.x;
    if (k >= m) return;
    const uint32_t a = cs[k].a, b = cs[k].b;
    meta[k] = make_uint4(a, b, rowFlags(bodies[a], bodies[b]), 0);
}

__global__ __launch_bounds__(kThreads) void kPrepareRows(const vx_body* __restrict__ bodies, vx_contact* __restrict__ cs,
                                                         uint32_t m, float h, SolveArrays sa) {
    uint32_t k = blockIdx.x * kThreads + threadIdx.x;
    if (k >= m) return;
    const vx_contact c = cs[k];
    const vx_body a = bodies[c.a], b = bodies[c.b];
    uint32_t flags = rowFlags(a, b);
    const uint32_t p = sa.pos[k];
    float4* row = sa.rows + 8 * (size_t)p;
    if (flags & kSensorRow) {
        sa.meta[p] = make_uint4(c.a, c.b, flags, 0);
        return;
    }
    const float4 fa = sa.frame[c.a], fb = sa.frame[c.b], ia = sa.invI[c.a], ib = sa.invI[c.b];
    const bool rotA = flags & kRotA, rotB = flags & kRotB;
    auto IA = [&](V3 v) { return frameMul(fa, ia, v, rotA); };
    auto IB = [&](V3 v) { return frameMul(fb, ib, v, rotB); };
    V3 pa = anchorWorld(a, v3(c.localAnchorA)), pb = anchorWorld(b, v3(c.localAnchorB));
    cs[k].point = raw((pa + pb) * 0.5f);
    V3 ra = pa - v3(a.position), rb = pb - v3(b.position);
    const V3 n = v3(c.normal);
    V3 raxn = cross(ra, n), rbxn = cross(rb, n);
    float kN = solverInvMass(a) + solverInvMass(b) + dot(raxn, IA(raxn)) + dot(rbxn, IB(rbxn));
    if (kN <= 0.0f) flags |= kSkip;
    float lg = c.bias0 + dot(n, pa - pb);
    float targetBase = lg > 0.0f ? -lg / h : 0.0f;
    V3 t1, t2;
    tangents(n, t1, t2);
    V3 raxt1 = cross(ra, t1), rbxt1 = cross(rb, t1);
    V3 raxt2 = cross(ra, t2), rbxt2 = cross(rb, t2);
    float k1 = solverInvMass(a) + solverInvMass(b) + dot(raxt1, IA(raxt1)) + dot(rbxt1, IB(rbxt1));
    float k2 = solverInvMass(a) + solverInvMass(b) + dot(raxt2, IA(raxt2)) + dot(rbxt2, IB(rbxt2));
    const float kN1 = dot(raxn, IA(raxt1)) + dot(rbxn, IB(rbxt1));
    const float kN2 = dot(raxn, IA(raxt2)) + dot(rbxn, IB(rbxt2));
    const float k12 = dot(raxt1, IA(raxt2)) + dot(rbxt1, IB(rbxt2));
    float kr1 = dot(t1, IA(t1)) + dot(t1, IB(t1));
    float kr2 = dot(t2, IA(t2)) + dot(t2, IB(t2));
    float ks = dot(n, IA(n)) + dot(n, IB(n));
    row[0] = make_float4(ra.x, ra.y, ra.z, kN);
    row[1] = make_float4(rb.x, rb.y, rb.z, k1);
    row[2] = make_float4(n.x, n.y, n.z, k2);
    row[3] = make_float4(t1.x, t1.y, t1.z, targetBase);
    row[4] = make_float4(t2.x, t2.y, t2.z, lg);
    row[5] = make_float4(c.vn0, c.restitution, c.friction1, c.friction2);
    row[6] = make_float4(c.rollingFriction, c.spinningFriction, kr1, kr2);
    row[7] = make_float4(ks, kN1, kN2, k12);
    sa.meta[p] = make_uint4(c.a, c.b, flags, 0);
    float4* sd = sa.side + 8 * (size_t)p;
    sd[0] = fa; sd[1] = ia; sd[2] = sa.vel[c.a]; sd[3] = sa.ang[c.a];
    sd[4] = fb; sd[5] = ib; sd[6] = sa.vel[c.b]; sd[7] = sa.ang[c.b];
}

__device__ inline V3 xyz(float4 v) { return {v.x, v.y, v.z}; }

// Velocity-dependent rows of solveContact (narrowphase.h:183-365) for one
// contact whose pose-dependent rows are already loaded.
struct RowRegs {
    float4 r[8];
    uint32_t flags;
};

template <bool kCone>
__device__ __attribute__((always_inline)) inline void solveRegs(const RowRegs& R, float4 fa, float4 iia, float4 fb, float4 iib, V3& vA, V3& wA, V3& vB,
                                 V3& wB, float& cN, float& cT1, float& cT2, float& cR1, float& cR2, float& cS) {
    const uint32_t flags = R.flags;
    const float4 r0 = R.r[0], r1 = R.r[1], r2 = R.r[2], r3 = R.r[3], r4 = R.r[4], r5 = R.r[5], r6 = R.r[6], r7 = R.r[7];
    const V3 ra = xyz(r0), rb = xyz(r1), n = xyz(r2), t1 = xyz(r3), t2 = xyz(r4);
    const float kN = r0.w, k1 = r1.w, k2 = r2.w, targetBase = r3.w, lg = r4.w;
    const float vn0 = r5.x, rest = r5.y, mu1 = r5.z, mu2 = r5.w;
    const float rolling = r6.x, spinning = r6.y, kr1 = r6.z, kr2 = r6.w;
    const float ks = r7.x, kN1 = r7.y, kN2 = r7.z, k12 = r7.w;
    const bool dynA = flags & kDynA, dynB = flags & kDynB, rotA = flags & kRotA, rotB = flags & kRotB;
    const float simA = iia.w, simB = iib.w;
    auto IA = [&](V3 v) { return frameMul(fa, iia, v, rotA); };
    auto IB = [&](V3 v) { return frameMul(fb, iib, v, rotB); };
    float vn = dot((vA + cross(wA, ra)) - (vB + cross(wB, rb)), n);
    float target = targetBase;
    if (vn0 < -1.0f && (lg <= 0.0f || vn > 0.0f)) target = fmx(target, -rest * vn0);
    if (!kCone) {
        float jn = (target - vn) / kN;
        float nextN = fmx(cN + jn, 0.0f);
        jn = nextN - cN;
        cN = nextN;
        V3 ni = n * jn;
        if (dynA) { vA += ni * simA; wA += IA(cross(ra, ni)); }
        if (dynB) { vB -= ni * simB; wB -= IB(cross(rb, ni)); }
        if (k1 > 0.0f && k2 > 0.0f) {
            V3 rv = (vA + cross(wA, ra)) - (vB + cross(wB, rb));
            float o1 = cT1, o2 = cT2;
            float n1 = o1 - dot(rv, t1) / k1;
            float n2 = o2 - dot(rv, t2) / k2;
            float m1 = mu1 * cN, m2 = mu2 * cN;
            if (m1 <= 0.0f) n1 = 0.0f;
            if (m2 <= 0.0f) n2 = 0.0f;
            float el = (m1 > 0.0f ? n1 * n1 / (m1 * m1) : 0.0f) + (m2 > 0.0f ? n2 * n2 / (m2 * m2) : 0.0f);
            if (el > 1.0f) { float sc = 1.0f / sqrtf(el); n1 *= sc; n2 *= sc; }
            cT1 = n1;
            cT2 = n2;
            V3 fi = t1 * (n1 - o1) + t2 * (n2 - o2);
            if (dynA) { vA += fi * simA; wA += IA(cross(ra, fi)); }
            if (dynB) { vB -= fi * simB; wB -= IB(cross(rb, fi)); }
        }
    } else {
        float on = cN, o1 = cT1, o2 = cT2;
        float nn = on, n1 = o1, n2 = o2;
        if (k1 > 0.0f && k2 > 0.0f) {
            const float det = kN * (k1 * k2 - k12 * k12) - kN1 * (kN1 * k2 - kN2 * k12) + kN2 * (kN1 * k12 - kN2 * k1);
            if (fabsf(det) > 1e-12f) {
                const V3 rv = (vA + cross(wA, ra)) - (vB + cross(wB, rb));
                const float bn = target - dot(rv, n);
                const float b1 = -dot(rv, t1);
                const float b2 = -dot(rv, t2);
                const float dn = (bn * (k1 * k2 - k12 * k12) + b1 * (kN2 * k12 - kN1 * k2) + b2 * (kN1 * k12 - kN2 * k1)) / det;
                const float d1 = (bn * (kN2 * k12 - kN1 * k2) + b1 * (kN * k2 - kN2 * kN2) + b2 * (kN1 * kN2 - kN * k12)) / det;
                const float d2 = (bn * (kN1 * k12 - kN2 * k1) + b1 * (kN1 * kN2 - kN * k12) + b2 * (kN * k1 - kN1 * kN1)) / det;
                nn = fmx(on + dn, 0.0f);
                n1 = o1 + d1;
                n2 = o2 + d2;
            }
        }
        const float m1 = mu1 * nn, m2 = mu2 * nn;
        if (m1 <= 0.0f) n1 = 0.0f;
        if (m2 <= 0.0f) n2 = 0.0f;
        const float el = (m1 > 0.0f ? n1 * n1 / (m1 * m1) : 0.0f) + (m2 > 0.0f ? n2 * n2 / (m2 * m2) : 0.0f);
        if (el > 1.0f) { const float sc = 1.0f / sqrtf(el); n1 *= sc; n2 *= sc; }
        cN = nn;
        cT1 = n1;
        cT2 = n2;
        const V3 imp = n * (nn - on) + t1 * (n1 - o1) + t2 * (n2 - o2);
        if (dynA) { vA += imp * simA; wA += IA(cross(ra, imp)); }
        if (dynB) { vB -= imp * simB; wB -= IB(cross(rb, imp)); }
    }
    V3 wr = wA - wB;
    if (rolling > 0.0f && kr1 > 0.0f && kr2 > 0.0f) {
        float o1 = cR1, o2 = cR2;
        float n1 = o1 - dot(wr, t1) / kr1;
        float n2 = o2 - dot(wr, t2) / kr2;
        float lim = rolling * cN;
        float mag = sqrtf(n1 * n1 + n2 * n2);
        if (mag > lim && mag > 1e-9f) { float sc = lim / mag; n1 *= sc; n2 *= sc; }
        cR1 = n1;
        cR2 = n2;
        V3 tq = t1 * (n1 - o1) + t2 * (n2 - o2);
        if (dynA) wA += IA(tq);
        if (dynB) wB -= IB(tq);
    }
    wr = wA - wB;
    if (spinning > 0.0f && ks > 0.0f) {
        float o = cS;
        float nx = o - dot(wr, n) / ks;
        float lim = spinning * cN;
        nx = fclamp(nx, -lim, lim);
        cS = nx;
        V3 tq = n * (nx - o);
        if (dynA) wA += IA(tq);
        if (dynB) wB -= IB(tq);
    }
}

// One contact on the compact global arrays.
template <bool kCone>
__device__ __attribute__((always_inline)) inline void solveRow(SolveArrays sa, uint32_t k) {
    const uint4 meta = sa.meta[k];
    if (meta.z & kSkip) return;
    const uint32_t ia = meta.x, ib = meta.y;
    RowRegs R;
    for (int q = 0; q < 8; ++q) R.r[q] = sa.rows[8 * (size_t)k + q];
    R.flags = meta.z;
    V3 vA = xyz(sa.vel[ia]), wA = xyz(sa.ang[ia]), vB = xyz(sa.vel[ib]), wB = xyz(sa.ang[ib]);
    float4 i0 = sa.imp[2 * k], i1 = sa.imp[2 * k + 1];
    float cN = i0.x, cT1 = i0.y, cT2 = i0.z, cR1 = i0.w, cR2 = i1.x, cS = i1.y;
    solveRegs<kCone>(R, sa.frame[ia], sa.invI[ia], sa.frame[ib], sa.invI[ib], vA, wA, vB, wB, cN, cT1, cT2, cR1, cR2, cS);
    if (meta.z & kDynA) { sa.vel[ia] = make_float4(vA.x, vA.y, vA.z, 0.0f); sa.ang[ia] = make_float4(wA.x, wA.y, wA.z, 0.0f); }
    if (meta.z & kDynB) { sa.vel[ib] = make_float4(vB.x, vB.y, vB.z, 0.0f); sa.ang[ib] = make_float4(wB.x, wB.y, wB.z, 0.0f); }
    sa.imp[2 * k] = make_float4(cN, cT1, cT2, cR1);
    sa.imp[2 * k + 1] = make_float4(cR2, cS, 0.0f, 0.0f);
}

}  // namespace vxk

namespace vxk {


// Workgroup barrier that orders LDS only: global loads in flight (the next
// rows' prefetch) and global stores (a row's own impulses, read back only by
// the same thread) are not waited for.
__device__ __attribute__((always_inline)) inline void ldsBarrier() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup", "local");
    __builtin_amdgcn_s_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup", "local");
}

// Runs of narrow big-island colours are solved in one workgroup, one colour
// after the other (a launch per colour costs more than a few rounds of 1024
// rows); the barrier orders the colours on one compute unit.  512 threads
// leave each wave the registers to hold rows loaded two passes ahead without
// spilling (kSweepColorRun, kSweepTailPair).
constexpr int kSweepRunThreads = 512;

// The tail of the big-island colours (the last colours, every one narrow)
// touches few bodies; their velocities then live in LDS for the whole sweep:
// loaded once, updated in place colour after colour, written back at the end.
// Two lanes per row as in the LDS batch sweeps (kSweepTailPair).  The tail is
// cut to kTailRows rows, so its slots always fit (the global-memory run in
// the kernel is a guard only), and to kTailColors colours, whose bounds are
// kept in LDS too.
constexpr uint32_t kTailCap = 6144;     // 6 floats x 6144 slots = 144 KB of velocities
constexpr uint32_t kTailRows = kTailCap / 2;
constexpr uint32_t kTailColors = 2047;  // + 1 bound x 4 B: 152 KB of LDS in all

// (body, 2 * row + side) entries of the tail rows' written bodies; the other
// sides sort last under `sentinel`.
__global__ __launch_bounds__(kThreads) void kTailEntries(SolveArrays sa, uint32_t r0, uint32_t rows, uint64_t sentinel,
                                                         uint64_t* __restrict__ keys, uint32_t* __restrict__ vals) {
    uint32_t i = blockIdx.x * kThreads + threadIdx.x;
    if (i >= rows) return;
    const uint4 m = sa.meta[r0 + i];
    keys[2 * i] = (m.z & kDynA) ? (uint64_t)m.x : sentinel;
    keys[2 * i + 1] = (m.z & kDynB) ? (uint64_t)m.y : sentinel;
    vals[2 * i] = 2 * i;
    vals[2 * i + 1] = 2 * i + 1;
}

__global__ __launch_bounds__(kThreads) void kTailHeads(const uint64_t* __restrict__ keys, uint32_t n, uint64_t sentinel,
                                                       uint32_t* __restrict__ head) {
    uint32_t i = blockIdx.x * kThreads + threadIdx.x;
    if (i >= n) return;
    head[i] = keys[i] != sentinel && (i == 0 || keys[i - 1] != keys[i]) ? 1u : 0u;
}

// Slots are 16-bit, side A in the low half of a row's 32-bit word and side B
// in the high half (the pair layout of the LDS sweeps); 0xffff for none.
__global__ __launch_bounds__(kThreads) void kTailSlots(const uint64_t* __restrict__ keys,
                                                       const uint32_t* __restrict__ vals, uint32_t n, uint64_t sentinel,
                                                       const uint32_t* __restrict__ head,
                                                       const uint32_t* __restrict__ headPos,
                                                       uint16_t* __restrict__ rowSlot, uint32_t* __restrict__ slotBody) {
    uint32_t i = blockIdx.x * kThreads + threadIdx.x;
    if (i >= n) return;
    if (keys[i] == sentinel) { rowSlot[vals[i]] = 0xffffu; return; }
    const uint32_t slot = headPos[i] + head[i] - 1;
    rowSlot[vals[i]] = (uint16_t)slot;
    if (head[i]) slotBody[slot] = (uint32_t)keys[i];
}

// --------------------------------------------------------------- islands --
// Connected components of the contact graph over dynamic bodies
// (the reference's buildIslands relation, solver.cpp:507-581).  Lock-free
// union-find: hook the larger root under the smaller with compare-and-swap;
// parent reads are device-coherent atomic loads.
__device__ inline uint32_t ufFind(uint32_t* parent, uint32_t x) {
    for (;;) {
        uint32_t p = __hip_atomic_load(&parent[x], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        if (p == x) return x;
        uint32_t gp = __hip_atomic_load(&parent[p], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        if (gp != p) __hip_atomic_store(&parent[x], gp, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        x = p;
    }
}

__global__ __launch_bounds__(kThreads) void kIslandUnion(const uint4* __restrict__ meta, uint32_t m, uint32_t* parent) {
    uint32_t k = blockIdx.x * kThreads + threadIdx.x;
    if (k >= m) return;
    const uint4 mt = meta[k];
    if ((mt.z & kSensorRow) || !(mt.z & kDynA) || !(mt.z & kDynB)) return;
    uint32_t a = mt.x, b = mt.y;
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

// Island of every contact (root of its dynamic body; kNone when it has none)
// and the number of contacts per island.
__global__ __launch_bounds__(kThreads) void kIslandOf(const uint4* __restrict__ meta, uint32_t m, uint32_t* parent,
                                                      uint32_t* __restrict__ islandOf, uint32_t* __restrict__ size) {
    uint32_t k = blockIdx.x * kThreads + threadIdx.x;
    uint32_t root = kNone;
    if (k < m) {
        const uint4 mt = meta[k];
        if (mt.z & kDynA) root = ufFind(parent, mt.x);
        else if (mt.z & kDynB) root = ufFind(parent, mt.y);
        islandOf[k] = root;
    }
    // The lanes in the first active lane's island add together (one atomic
    // for a wavefront inside one big island); the others add their own.
    const uint64_t act = __ballot(root != kNone);
    if (!act) return;
    const uint32_t lead = (uint32_t)__builtin_amdgcn_readlane((int)root, __ffsll((unsigned long long)act) - 1);
    const uint64_t same = __ballot(root == lead);
    if (root == lead) {
        if ((threadIdx.x & 63) == (uint32_t)(__ffsll((unsigned long long)same) - 1))
            atomicAdd(&size[lead], (uint32_t)__popcll(same));
    } else if (root != kNone) {
        atomicAdd(&size[root], 1u);
    }
}

// Islands of up to kLdsIsland contacts are packed, in root order, into
// batches of about kBatch contacts (an island joins the batch its first
// contact's offset falls in) whose dynamic bodies live in LDS while the batch
// is solved.  A batch holds fewer than kBatch + kLdsIsland contacts and at
// most kBatch islands, and an island of c contacts has at most c + 1 bodies,
// so kLdsBodies slots always suffice.  Islands up to the big-island limit
// get a batch of their own on global memory; bigger ones go colour by colour
// over the whole device.
constexpr uint32_t kBatch = 512;
constexpr uint32_t kLdsIsland = 2048;
constexpr uint32_t kLdsBodies = kBatch + kLdsIsland + kBatch;

// Per island root: contact count if the island goes through LDS batches (0
// otherwise) and a flag for the global-memory islands; counts islands.
__global__ __launch_bounds__(kThreads) void kIslandKinds(const uint32_t* __restrict__ size, uint32_t n, uint32_t limit,
                                                         uint32_t* __restrict__ ldsSize, uint32_t* __restrict__ globalFlag,
                                                         uint32_t* __restrict__ islands, uint32_t* __restrict__ largest) {
    uint32_t i = blockIdx.x * kThreads + threadIdx.x;
    const uint32_t sz = i < n ? size[i] : 0u;
    if (sz) atomicMax(largest, sz);
    const bool lds = sz > 0 && sz <= kLdsIsland && sz <= limit;
    const bool global = sz > kLdsIsland && sz <= limit;
    if (i < n) {
        ldsSize[i] = lds ? sz : 0u;
        globalFlag[i] = global ? 1u : 0u;
    }
    waveReserve(lds || global ? 1u : 0u, islands);
}

// Splits contacts (in colour order) into the batched list, keyed by
// (batch, contact) so that a batch lists its contacts in colour order, and
// the big-island list (colour order kept).  LDS batches take ids below
// globalBase; global-memory islands follow, one batch each.
__global__ __launch_bounds__(kThreads) void kIslandClassify(const uint32_t* __restrict__ islandOf,
                                                            const uint32_t* __restrict__ size,
                                                            const uint32_t* __restrict__ ldsOffset,
                                                            const uint32_t* __restrict__ globalRank, uint32_t globalBase,
                                                            uint32_t m, uint32_t limit, uint32_t* __restrict__ bigFlag,
                                                            uint64_t* __restrict__ smallKeys,
                                                            uint32_t* __restrict__ smallVals,
                                                            uint32_t* __restrict__ smallCounter, int mb) {
    uint32_t k = blockIdx.x * kThreads + threadIdx.x;
    uint32_t root = k < m ? islandOf[k] : kNone;
    const uint32_t sz = root != kNone ? size[root] : 0u;
    const bool big = sz > limit;
    const bool small = root != kNone && !big;
    if (k < m) bigFlag[k] = big ? 1u : 0u;
    uint32_t at = waveReserve(small ? 1u : 0u, smallCounter);
    if (small) {
        const uint64_t batch = sz <= kLdsIsland ? ldsOffset[root] / kBatch : globalBase + globalRank[root];
        smallKeys[at] = (batch << mb) | k;
        smallVals[at] = k;
    }
}

// Dynamic bodies of LDS-batched islands: rank within the island.
__global__ __launch_bounds__(kThreads) void kBatchBodyCount(const vx_body* __restrict__ bodies, uint32_t n,
                                                            uint32_t* parent, const uint32_t* __restrict__ ldsSize,
                                                            uint32_t* __restrict__ bodyRoot,
                                                            uint32_t* __restrict__ bodyRank,
                                                            uint32_t* __restrict__ bodyCount) {
    uint32_t i = blockIdx.x * kThreads + threadIdx.x;
    if (i >= n) return;
    uint32_t root = kNone;
    if (isDynamic(bodies[i])) {
        root = ufFind(parent, i);
        if (ldsSize[root] == 0) root = kNone;
    }
    bodyRoot[i] = root;
    if (root != kNone) bodyRank[i] = atomicAdd(&bodyCount[root], 1u);
}

// Body range of each LDS batch (indexed by batch id).
__global__ __launch_bounds__(kThreads) void kBatchBodyRange(const uint32_t* __restrict__ bodyCount,
                                                            const uint32_t* __restrict__ bodyOffset,
                                                            const uint32_t* __restrict__ ldsOffset, uint32_t n,
                                                            uint32_t* __restrict__ rangeStart,
                                                            uint32_t* __restrict__ rangeEnd) {
    uint32_t r = blockIdx.x * kThreads + threadIdx.x;
    const bool has = r < n && bodyCount[r] != 0;
    const uint32_t b = has ? ldsOffset[r] / kBatch : kNone;
    uint32_t lo = has ? bodyOffset[r] : 0xffffffffu, hi = has ? bodyOffset[r] + bodyCount[r] : 0u;
    // Roots are in batch order, so a wavefront's roots mostly share the first
    // one's batch: those combine into one atomic pair, the others go alone.
    const uint64_t act = __ballot(has);
    if (!act) return;
    const uint32_t lead = (uint32_t)__builtin_amdgcn_readlane((int)b, __ffsll((unsigned long long)act) - 1);
    const bool same = has && b == lead;
    uint32_t slo = same ? lo : 0xffffffffu, shi = same ? hi : 0u;
    for (int off = 32; off > 0; off >>= 1) {
        slo = min(slo, (uint32_t)__shfl_xor((int)slo, off));
        shi = max(shi, (uint32_t)__shfl_xor((int)shi, off));
    }
    if ((threadIdx.x & 63) == 0) {
        atomicMin(&rangeStart[lead], slo);
        atomicMax(&rangeEnd[lead], shi);
    }
    if (has && !same) {
        atomicMin(&rangeStart[b], lo);
        atomicMax(&rangeEnd[b], hi);
    }
}

// Body list of the batches and each body's LDS slot within its batch.
__global__ __launch_bounds__(kThreads) void kBatchBodySlots(const uint32_t* __restrict__ bodyRoot,
                                                            const uint32_t* __restrict__ bodyRank,
                                                            const uint32_t* __restrict__ bodyOffset,
                                                            const uint32_t* __restrict__ ldsOffset,
                                                            const uint32_t* __restrict__ rangeStart, uint32_t n,
                                                            uint32_t* __restrict__ bodyList, uint32_t* __restrict__ slotOf) {
    uint32_t i = blockIdx.x * kThreads + threadIdx.x;
    if (i >= n) return;
    const uint32_t root = bodyRoot[i];
    if (root == kNone) return;
    const uint32_t pos = bodyOffset[root] + bodyRank[i];
    bodyList[pos] = i;
    slotOf[i] = pos - rangeStart[ldsOffset[root] / kBatch];
}

// LDS slots of both endpoints of every batched contact (kNone's low half
// where the endpoint is not a dynamic body of the batch).
__global__ __launch_bounds__(kThreads) void kBatchContactSlots(const uint64_t* __restrict__ keys,
                                                               const uint32_t* __restrict__ idx, uint32_t ms,
                                                               const uint4* __restrict__ meta,
                                                               const uint32_t* __restrict__ slotOf, uint32_t globalBase,
                                                               int mb, uint32_t* __restrict__ slotPair) {
    uint32_t j = blockIdx.x * kThreads + threadIdx.x;
    if (j >= ms) return;
    uint32_t sa = 0xffffu, sb = 0xffffu;
    if ((keys[j] >> mb) < globalBase) {
        const uint4 mt = meta[idx[j]];
        if (!(mt.z & kSensorRow)) {
            if (mt.z & kDynA) sa = slotOf[mt.x];
            if (mt.z & kDynB) sb = slotOf[mt.y];
        }
    }
    slotPair[j] = sa | (sb << 16);
}

// Colour counts go through an LDS histogram when the K colours fit it: one
// global atomic per colour and workgroup instead of one per contact.
constexpr uint32_t kBigHist = 2048;
__global__ __launch_bounds__(kThreads) void kBigList(const uint32_t* __restrict__ bigFlag,
                                                     const uint32_t* __restrict__ bigPos, uint32_t m,
                                                     const uint32_t* __restrict__ color, uint32_t* __restrict__ bigIdx,
                                                     uint32_t* __restrict__ bigColorCount, uint32_t K) {
    __shared__ uint32_t hist[kBigHist];
    const bool lds = K <= kBigHist;
    if (lds)
        for (uint32_t q = threadIdx.x; q < K; q += kThreads) hist[q] = 0;
    __syncthreads();
    const uint32_t k = blockIdx.x * kThreads + threadIdx.x;
    if (k < m && bigFlag[k]) {
        bigIdx[bigPos[k]] = k;
        if (lds) atomicAdd(&hist[color[k]], 1u);
        else atomicAdd(&bigColorCount[color[k]], 1u);
    }
    if (!lds) return;
    __syncthreads();
    for (uint32_t q = threadIdx.x; q < K; q += kThreads)
        if (hist[q]) atomicAdd(&bigColorCount[q], hist[q]);
}

// Solve position of every contact: batched contacts at their place in the
// batch-sorted list, big-island contacts after them in colour order.
__global__ __launch_bounds__(kThreads) void kSolvePositions(const uint32_t* __restrict__ idx, uint32_t ms,
                                                            const uint32_t* __restrict__ bigIdx, uint32_t mb,
                                                            uint32_t* __restrict__ pos) {
    uint32_t j = blockIdx.x * kThreads + threadIdx.x;
    if (j < ms) pos[idx[j]] = j;
    else if (j < ms + mb) pos[bigIdx[j - ms]] = j;
}

// Run heads of the batch-sorted list: a run is a maximal stretch of one
// batch and one colour.  Counts the LDS batches.
__global__ __launch_bounds__(kThreads) void kRunHeads(const uint64_t* __restrict__ keys, const uint32_t* __restrict__ idx,
                                                      uint32_t ms, const uint32_t* __restrict__ color,
                                                      uint32_t* __restrict__ runHead, uint32_t* __restrict__ batchHead,
                                                      int mb, uint32_t globalBase, uint32_t* __restrict__ ldsBatches) {
    uint32_t j = blockIdx.x * kThreads + threadIdx.x;
    if (j >= ms) return;
    bool newBatch = j == 0 || (keys[j] >> mb) != (keys[j - 1] >> mb);
    bool newRun = newBatch || color[idx[j]] != color[idx[j - 1]];
    runHead[j] = newRun ? 1u : 0u;
    batchHead[j] = newBatch ? 1u : 0u;
    if (newBatch && (keys[j] >> mb) < globalBase) atomicAdd(ldsBatches, 1u);
}

__global__ __launch_bounds__(kThreads) void kRunTables(const uint64_t* __restrict__ keys,
                                                       const uint32_t* __restrict__ runHead,
                                                       const uint32_t* __restrict__ batchHead,
                                                       const uint32_t* __restrict__ runPos,
                                                       const uint32_t* __restrict__ batchPos, uint32_t ms, int mb,
                                                       uint32_t globalBase, const uint32_t* __restrict__ rangeStart,
                                                       const uint32_t* __restrict__ rangeEnd,
                                                       uint32_t* __restrict__ runStart,
                                                       uint32_t* __restrict__ batchFirstRun,
                                                       uint2* __restrict__ batchBodies) {
    uint32_t j = blockIdx.x * kThreads + threadIdx.x;
    if (j >= ms) return;
    if (runHead[j]) runStart[runPos[j]] = j;
    if (batchHead[j]) {
        batchFirstRun[batchPos[j]] = runPos[j];
        const uint32_t b = (uint32_t)(keys[j] >> mb);
        batchBodies[batchPos[j]] = b < globalBase ? make_uint2(rangeStart[b], rangeEnd[b]) : make_uint2(0u, 0u);
    }
}

// Solver passes per sweep of each batch (a run of L contacts takes
// ceil(L / (kThreads/2)) passes); the maximum over batches.
__global__ __launch_bounds__(kThreads) void kBatchPasses(const uint32_t* __restrict__ runStart,
                                                         const uint32_t* __restrict__ batchFirstRun,
                                                         const uint32_t* __restrict__ batches,
                                                         uint32_t* __restrict__ maxPasses,
                                                         const uint2* __restrict__ batchBodies,
                                                         uint32_t* __restrict__ maxBodies, uint32_t* __restrict__ maxRows) {
    uint32_t q = blockIdx.x * kThreads + threadIdx.x;
    if (q >= *batches) return;
    uint32_t passes = 0;
    for (uint32_t r = batchFirstRun[q]; r < batchFirstRun[q + 1]; ++r)
        passes += (runStart[r + 1] - runStart[r] + kThreads / 2 - 1) / (kThreads / 2);
    atomicMax(maxPasses, passes);
    atomicMax(maxBodies, batchBodies[q].y - batchBodies[q].x);
    atomicMax(maxRows, runStart[batchFirstRun[q + 1]] - runStart[batchFirstRun[q]]);
}

__global__ void kRunSentinels(uint32_t* runStart, const uint32_t* runs, uint32_t ms, uint32_t* batchFirstRun,
                              const uint32_t* batches, const uint32_t* ldsBatches, uint32_t* ranges,
                              uint32_t* ldsContacts) {
    runStart[*runs] = ms;
    batchFirstRun[*batches] = *runs;
    ranges[0] = 0; ranges[1] = *ldsBatches;         // LDS batches
    ranges[2] = *ldsBatches; ranges[3] = *batches;  // global-memory islands
    *ldsContacts = runStart[batchFirstRun[*ldsBatches]];
}


// ------------------------------------------------------ LDS batch sweeps --
// Two lanes per contact: the even lane carries body A, the odd lane body B.
// Each lane keeps its own body's velocity update and inertia products; the
// quantities both sides need (relative velocity, impulses) are formed from
// the same operands in the same order on both lanes after a lane swap, so
// the two copies agree bit for bit and equal the one-lane sequence of
// solveContact (narrowphase.h:183-365).  b -= x is computed as b + (-x),
// which IEEE defines to be the same operation.

// Every lane reads a lane of its own quad, so no lane keeps its old value:
// the move takes no initial value (no register to clear first).
__device__ inline float swapLane(float x) {  // lanes 2i <-> 2i+1
    return __int_as_float(__builtin_amdgcn_mov_dpp(__float_as_int(x), 0xB1, 0xF, 0xF, true));
}
__device__ inline V3 swapLane(V3 v) { return {swapLane(v.x), swapLane(v.y), swapLane(v.z)}; }

// What one lane needs of one contact, all addressed by its solve position
// (loaded one pass ahead of its use).
struct PairItem {
    uint4 meta;
    uint32_t slots;
    float4 r[8];
    float4 frame, invI, vel, ang;  // this lane's body
    float4 i0, i1;                 // accumulated impulses
};

// kImp: also load the accumulated impulses (kernels that keep them in LDS
// read them there at the row's turn instead).
template <bool kImp = true>
__device__ __attribute__((always_inline)) inline PairItem loadPairItemSlots(SolveArrays sa, uint32_t slots, uint32_t j,
                                                                           bool sideB) {
    PairItem it;
    it.meta = sa.meta[j];
    it.slots = slots;
    const float4* row = sa.rows + 8 * (size_t)j;
#pragma unroll
    for (int q = 0; q < 8; ++q) it.r[q] = row[q];
    const float4* sd = sa.side + 8 * (size_t)j + (sideB ? 4 : 0);
    it.frame = sd[0]; it.invI = sd[1]; it.vel = sd[2]; it.ang = sd[3];
    if (kImp) {
        it.i0 = sa.imp[2 * j];
        it.i1 = sa.imp[2 * j + 1];
    }
    return it;
}

template <bool kImp = true>
__device__ __attribute__((always_inline)) inline PairItem loadPairItem(SolveArrays sa, const uint32_t* slotPair,
                                                                      uint32_t j, bool sideB) {
    return loadPairItemSlots<kImp>(sa, slotPair[j], j, sideB);
}

// The row data one lane of an LDS-resident pass reads, loaded at exactly the
// widths it uses (the flags word of meta, the arm, direction and mass terms
// the mode reads, the side's frame, inertia and xyz velocities): a loaded
// register that no instruction reads may be reused while its load is still in
// flight, and the wait for that load would stall the pass on the loads of the
// next one.  The impulses are read from LDS at the row's turn.
template <int kMode>
__device__ __attribute__((always_inline)) inline PairItem loadLdsRow(SolveArrays sa, uint32_t slots, uint32_t j,
                                                                    bool sideB) {
    PairItem it;
    it.meta = make_uint4(0u, 0u, reinterpret_cast<const uint32_t*>(sa.meta + j)[2], 0u);
    it.slots = slots;
    const float4* row = sa.rows + 8 * (size_t)j;
    if (kMode == 2) {  // warm start: arms and the three directions
#pragma unroll
        for (int q = 0; q < 5; ++q) {
            const float* f = reinterpret_cast<const float*>(row + q);
            it.r[q] = make_float4(f[0], f[1], f[2], 0.0f);
        }
#pragma unroll
        for (int q = 5; q < 8; ++q) it.r[q] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    } else {
#pragma unroll
        for (int q = 0; q < 7; ++q) it.r[q] = row[q];
        if (kMode == 1) it.r[7] = row[7];
        else it.r[7] = make_float4(reinterpret_cast<const float*>(row + 7)[0], 0.0f, 0.0f, 0.0f);  // ks only
    }
    const float4* sd = sa.side + 8 * (size_t)j + (sideB ? 4 : 0);
    it.frame = sd[0];
    it.invI = sd[1];
    const float* v = reinterpret_cast<const float*>(sd + 2);
    const float* w = reinterpret_cast<const float*>(sd + 3);
    it.vel = make_float4(v[0], v[1], v[2], 0.0f);
    it.ang = make_float4(w[0], w[1], w[2], 0.0f);
    return it;
}

// kGlobal: the lane's body velocities are read from and written back to the
// global arrays (slot and lv unused) instead of LDS slots.  kStoreImp: the
// new impulses are stored to the global array; without it they are only
// left in it.i0/it.i1, and the return value says whether they changed
// (a solved row, mode 0 or 1) so the caller can keep them elsewhere.
template <int kMode, uint32_t kStride = kLdsBodies, bool kGlobal = false, bool kStoreImp = true>
__device__ __attribute__((always_inline)) inline bool pairRow(SolveArrays sa, uint32_t j, PairItem& it, bool sideB,
                                                              float* __restrict__ lv) {
    const uint32_t flags = it.meta.z;
    if (kMode == 2 ? (flags & kSensorRow) != 0 : (flags & kSkip) != 0) return false;
    const bool dyn = flags & (sideB ? kDynB : kDynA);
    const bool rot = flags & (sideB ? kRotB : kRotA);
    const uint32_t slot = sideB ? it.slots >> 16 : it.slots & 0xffffu;
    const uint32_t body = sideB ? it.meta.y : it.meta.x;
    const float4 r0 = it.r[0], r1 = it.r[1], r2 = it.r[2], r3 = it.r[3], r4 = it.r[4], r5 = it.r[5], r6 = it.r[6],
                 r7 = it.r[7];
    const float4 fr = it.frame, ii = it.invI;
    V3 v, w;
    if (dyn && kGlobal) {
        v = xyz(sa.vel[body]);
        w = xyz(sa.ang[body]);
    } else if (dyn) {
        v = V3{lv[slot], lv[kStride + slot], lv[2 * kStride + slot]};
        w = V3{lv[3 * kStride + slot], lv[4 * kStride + slot], lv[5 * kStride + slot]};
    } else {
        v = xyz(it.vel);
        w = xyz(it.ang);
    }
    const V3 r = sideB ? xyz(r1) : xyz(r0);
    const V3 n = xyz(r2), t1 = xyz(r3), t2 = xyz(r4);
    const float sim = ii.w;
    auto I = [&](V3 x) { return frameMul(fr, ii, x, rot); };
    // Adds (side A) or subtracts (side B) a linear and an angular change.
    auto push = [&](V3 imp) {
        if (!dyn) return;
        const V3 dv = imp * sim;
        const V3 dw = I(cross(r, imp));
        v = v + (sideB ? V3{-dv.x, -dv.y, -dv.z} : dv);
        w = w + (sideB ? V3{-dw.x, -dw.y, -dw.z} : dw);
    };
    auto pushAng = [&](V3 tq) {
        if (!dyn) return;
        const V3 dw = I(tq);
        w = w + (sideB ? V3{-dw.x, -dw.y, -dw.z} : dw);
    };
    float cN = it.i0.x, cT1 = it.i0.y, cT2 = it.i0.z, cR1 = it.i0.w, cR2 = it.i1.x, cS = it.i1.y;
    if (kMode == 2) {
        if (!(cN <= 0.0f && cT1 == 0.0f && cT2 == 0.0f && cR1 == 0.0f && cR2 == 0.0f && cS == 0.0f) && dyn) {
            const V3 imp = n * cN + t1 * cT1 + t2 * cT2;
            const V3 ang = t1 * cR1 + t2 * cR2 + n * cS;
            const V3 dv = imp * sim;
            v = v + (sideB ? V3{-dv.x, -dv.y, -dv.z} : dv);
            const V3 dw1 = I(cross(r, imp));
            w = w + (sideB ? V3{-dw1.x, -dw1.y, -dw1.z} : dw1);
            const V3 dw2 = I(ang);
            w = w + (sideB ? V3{-dw2.x, -dw2.y, -dw2.z} : dw2);
        }
    } else {
        const float kN = r0.w, k1 = r1.w, k2 = r2.w, targetBase = r3.w, lg = r4.w;
        const float vn0 = r5.x, rest = r5.y, mu1 = r5.z, mu2 = r5.w;
        const float rolling = r6.x, spinning = r6.y, kr1 = r6.z, kr2 = r6.w;
        const float ks = r7.x, kN1 = r7.y, kN2 = r7.z, k12 = r7.w;
        // Relative point velocity (A minus B), identical on both lanes.
        auto relVel = [&]() {
            const V3 mine = v + cross(w, r);
            const V3 other = swapLane(mine);
            return sideB ? other - mine : mine - other;
        };
        auto relAng = [&]() {
            const V3 other = swapLane(w);
            return sideB ? other - w : w - other;
        };
        const float vn = dot(relVel(), n);
        float target = targetBase;
        if (vn0 < -1.0f && (lg <= 0.0f || vn > 0.0f)) target = fmx(target, -rest * vn0);
        if (kMode == 0) {
            float jn = (target - vn) / kN;
            const float nextN = fmx(cN + jn, 0.0f);
            jn = nextN - cN;
            cN = nextN;
            push(n * jn);
            if (k1 > 0.0f && k2 > 0.0f) {
                const V3 rv = relVel();
                const float o1 = cT1, o2 = cT2;
                float n1 = o1 - dot(rv, t1) / k1;
                float n2 = o2 - dot(rv, t2) / k2;
                const float m1 = mu1 * cN, m2 = mu2 * cN;
                if (m1 <= 0.0f) n1 = 0.0f;
                if (m2 <= 0.0f) n2 = 0.0f;
                const float el = (m1 > 0.0f ? n1 * n1 / (m1 * m1) : 0.0f) + (m2 > 0.0f ? n2 * n2 / (m2 * m2) : 0.0f);
                if (el > 1.0f) { const float sc = 1.0f / sqrtf(el); n1 *= sc; n2 *= sc; }
                cT1 = n1;
                cT2 = n2;
                push(t1 * (n1 - o1) + t2 * (n2 - o2));
            }
        } else {
            const float on = cN, o1 = cT1, o2 = cT2;
            float nn = on, n1 = o1, n2 = o2;
            if (k1 > 0.0f && k2 > 0.0f) {
                const float det = kN * (k1 * k2 - k12 * k12) - kN1 * (kN1 * k2 - kN2 * k12) + kN2 * (kN1 * k12 - kN2 * k1);
                if (fabsf(det) > 1e-12f) {
                    const V3 rv = relVel();
                    const float bn = target - dot(rv, n);
                    const float b1 = -dot(rv, t1);
                    const float b2 = -dot(rv, t2);
                    const float dn = (bn * (k1 * k2 - k12 * k12) + b1 * (kN2 * k12 - kN1 * k2) + b2 * (kN1 * k12 - kN2 * k1)) / det;
                    const float d1 = (bn * (kN2 * k12 - kN1 * k2) + b1 * (kN * k2 - kN2 * kN2) + b2 * (kN1 * kN2 - kN * k12)) / det;
                    const float d2 = (bn * (kN1 * k12 - kN2 * k1) + b1 * (kN1 * kN2 - kN * k12) + b2 * (kN * k1 - kN1 * kN1)) / det;
                    nn = fmx(on + dn, 0.0f);
                    n1 = o1 + d1;
                    n2 = o2 + d2;
                }
            }
            const float m1 = mu1 * nn, m2 = mu2 * nn;
            if (m1 <= 0.0f) n1 = 0.0f;
            if (m2 <= 0.0f) n2 = 0.0f;
            const float el = (m1 > 0.0f ? n1 * n1 / (m1 * m1) : 0.0f) + (m2 > 0.0f ? n2 * n2 / (m2 * m2) : 0.0f);
            if (el > 1.0f) { const float sc = 1.0f / sqrtf(el); n1 *= sc; n2 *= sc; }
            cN = nn;
            cT1 = n1;
            cT2 = n2;
            push(n * (nn - on) + t1 * (n1 - o1) + t2 * (n2 - o2));
        }
        if (rolling > 0.0f && kr1 > 0.0f && kr2 > 0.0f) {
            const V3 wr = relAng();
            const float o1 = cR1, o2 = cR2;
            float n1 = o1 - dot(wr, t1) / kr1;
            float n2 = o2 - dot(wr, t2) / kr2;
            const float lim = rolling * cN;
            const float mag = sqrtf(n1 * n1 + n2 * n2);
            if (mag > lim && mag > 1e-9f) { const float sc = lim / mag; n1 *= sc; n2 *= sc; }
            cR1 = n1;
            cR2 = n2;
            pushAng(t1 * (n1 - o1) + t2 * (n2 - o2));
        }
        if (spinning > 0.0f && ks > 0.0f) {
            const V3 wr = relAng();
            const float o = cS;
            float nx = o - dot(wr, n) / ks;
            const float lim = spinning * cN;
            nx = fclamp(nx, -lim, lim);
            cS = nx;
            pushAng(n * (nx - o));
        }
        // Both lanes store the (identical) impulses, so each lane's later
        // reload of this contact follows its own store.
        it.i0 = make_float4(cN, cT1, cT2, cR1);
        it.i1 = make_float4(cR2, cS, 0.0f, 0.0f);
        if (kStoreImp) {
            sa.imp[2 * j] = it.i0;
            sa.imp[2 * j + 1] = it.i1;
        }
    }
    if (dyn && kGlobal) {
        sa.vel[body] = make_float4(v.x, v.y, v.z, 0.0f);
        sa.ang[body] = make_float4(w.x, w.y, w.z, 0.0f);
    } else if (dyn) {
        lv[slot] = v.x; lv[kStride + slot] = v.y; lv[2 * kStride + slot] = v.z;
        lv[3 * kStride + slot] = w.x; lv[4 * kStride + slot] = w.y; lv[5 * kStride + slot] = w.z;
    }
    return kMode != 2;
}

// Medium islands, one workgroup each on the global velocity arrays: every
// sweep of every colour, two lanes per row, barrier between colours (the
// island shares no dynamic body with any other contact, and rows of one
// colour share none with each other, so this is the colour order exactly).
// Islands [first, last) of the island list.
template <int kMode, int kBlock>
__global__ __launch_bounds__(kBlock) void kIslandSweeps(SolveArrays sa, const uint32_t* __restrict__ runStart,
                                                        const uint32_t* __restrict__ islandFirstRun,
                                                        const uint32_t* __restrict__ islandRange, int sweeps) {
    const uint32_t first = islandRange[0], last = islandRange[1];
    const bool sideB = threadIdx.x & 1;
    const uint32_t half = threadIdx.x >> 1;
    constexpr uint32_t kPass = kBlock / 2;
    for (uint32_t isl = first + blockIdx.x; isl < last; isl += gridDim.x) {
        const uint32_t r0 = islandFirstRun[isl], r1 = islandFirstRun[isl + 1];
        for (int s = 0; s < sweeps; ++s)
            for (uint32_t r = r0; r < r1; ++r) {
                const uint32_t j0 = runStart[r], j1 = runStart[r + 1];
                for (uint32_t k = j0 + half; k < j1; k += kPass) {  // both lanes of a row together
                    PairItem it = loadPairItemSlots(sa, 0u, k, sideB);
                    pairRow<kMode, kLdsBodies, true>(sa, k, it, sideB, nullptr);
                }
                __syncthreads();
            }
    }
}

// A run of narrow colours c0..c1-1 by one workgroup, two lanes per row on
// the global velocity arrays, barrier between colours; rows are loaded two
// passes ahead (the tail kernel's cursor, with the colour bounds read from
// global memory).
template <int kMode>  // 0 two-axis, 1 cone-block, 2 warm start
__global__ __launch_bounds__(kSweepRunThreads) void kSweepColorRun(SolveArrays sa, uint32_t base,
                                                                   const uint32_t* __restrict__ colorStart, uint32_t c0,
                                                                   uint32_t c1) {
    const uint32_t t = threadIdx.x;
    const bool sideB = t & 1;
    const uint32_t half = t >> 1;
    constexpr uint32_t kPass = kSweepRunThreads / 2;
    struct Cursor { uint32_t c, cb, ce; };
    auto next = [&](const Cursor& x) {
        if (x.cb + kPass < x.ce) return Cursor{x.c, x.cb + kPass, x.ce};
        const uint32_t nc = x.c + 1;
        return nc < c1 ? Cursor{nc, colorStart[nc], colorStart[nc + 1]} : Cursor{c1, 0u, 0u};
    };
    Cursor A{c0, colorStart[c0], colorStart[c0 + 1]};
    PairItem itA, itB;
    if (A.cb + half < A.ce) itA = loadPairItemSlots(sa, 0u, base + A.cb + half, sideB);
    Cursor B = next(A);
    if (B.c < c1 && B.cb + half < B.ce) itB = loadPairItemSlots(sa, 0u, base + B.cb + half, sideB);
    for (;;) {
        const Cursor C = B.c < c1 ? next(B) : Cursor{c1, 0u, 0u};
        PairItem itC;
        if (C.c < c1 && C.cb + half < C.ce) itC = loadPairItemSlots(sa, 0u, base + C.cb + half, sideB);
        if (A.cb + half < A.ce) pairRow<kMode, kLdsBodies, true>(sa, base + A.cb + half, itA, sideB, nullptr);
        if (A.cb + kPass >= A.ce) __syncthreads();
        if (B.c >= c1) break;
        A = B; itA = itB;
        B = C; itB = itC;
    }
}

// One wide colour over the whole device, two lanes per row (the colour's
// rows share no body, so every row is independent), velocities in the global
// arrays.
template <int kMode>  // 0 two-axis, 1 cone-block, 2 warm start
__global__ __launch_bounds__(kThreads) void kSweepColorPair(SolveArrays sa, uint32_t base, uint32_t first,
                                                            uint32_t count) {
    const uint32_t g = blockIdx.x * kThreads + threadIdx.x;
    const uint32_t row = g >> 1;
    if (row >= count) return;  // both lanes of a row leave together
    const bool sideB = g & 1;
    const uint32_t k = base + first + row;
    PairItem it = loadPairItemSlots(sa, 0u, k, sideB);
    pairRow<kMode, kLdsBodies, true>(sa, k, it, sideB, nullptr);
}

// LDS batches, one workgroup each: the batch's dynamic bodies and its
// contacts' accumulated impulses are loaded into LDS, then every sweep of
// every colour present in the batch runs with a barrier between colours (the
// batch's islands share no dynamic body with any other contact, and contacts
// of one colour share none with each other, so this is the colour order
// exactly), and bodies and impulses are written back.  A pass is up to
// kThreads/2 contacts of one colour; each lane loads its next pass's row data
// while it solves the current one.  The sweeps make no global store and
// every lane issues the same loads on every pass (lanes without a row in the
// next pass load the batch's first row and drop it), so the wait for a
// pass's data never has to wait for the loads of the pass after it.
constexpr uint32_t kBatchRows = kBatch + kLdsIsland;  // more than any batch's contact count

// Persistent form: a grid of at most one workgroup per CU (the batch's LDS
// allows no more; three with kSmall, whose LDS holds batches of up to
// kSmallBodies bodies and kSmallRows contacts) takes batches from a device
// queue, in the order `order`
// lists them (longest first: the batches' pass counts, so the long chains
// start early and the short ones fill in behind them), or in batch order
// without one.  queue[0] hands out slots, queue[1] counts finished
// workgroups; the last one to finish clears both for the next launch.
constexpr uint32_t kSmallBodies = 1024, kSmallRows = 1024;

template <int kMode, int kBlock = kThreads, bool kSmall = false>
__global__ __launch_bounds__(kBlock) void kBatchSweeps(SolveArrays sa, const uint32_t* __restrict__ slotPair,
                                                       const uint32_t* __restrict__ runStart,
                                                       const uint32_t* __restrict__ batchFirstRun,
                                                       const uint2* __restrict__ batchBodies,
                                                       const uint32_t* __restrict__ bodyList,
                                                       const uint32_t* __restrict__ batchRange,
                                                       const uint32_t* __restrict__ order, uint32_t* queue, int sweeps,
                                                       uint32_t* __restrict__ error) {
    constexpr uint32_t kBodies = kSmall ? kSmallBodies : kLdsBodies, kRows = kSmall ? kSmallRows : kBatchRows;
    __shared__ float lv[6 * kBodies];
    __shared__ float4 li4[kRows];  // cN, cT1, cT2, cR1
    __shared__ float2 li2[kRows];  // cR2, cS
    __shared__ uint32_t slotS;
    const uint32_t first = batchRange[0], last = batchRange[1];
    const uint32_t t = threadIdx.x;
    const bool sideB = t & 1;
    constexpr uint32_t kPass = kBlock / 2;
    for (;;) {
        if (t == 0) slotS = atomicAdd(&queue[0], 1u);
        __syncthreads();
        const uint32_t slot = slotS;
        if (slot >= last - first) break;
        const uint32_t q = first + (order ? order[slot] : slot);
        const uint2 bb = batchBodies[q];
        const uint32_t nb = bb.y - bb.x;
        const uint32_t r0 = batchFirstRun[q], r1 = batchFirstRun[q + 1];
        const uint32_t jb = runStart[r0], nc = runStart[r1] - jb;  // the batch's contacts, contiguous
        if (nb > kBodies || nc > kRows) {  // excluded by the packing bounds; reported, never overrun
            if (t == 0) atomicOr(error, 1u);
            __syncthreads();
            continue;
        }
        for (uint32_t i = t; i < nb; i += kBlock) {
            const uint32_t body = bodyList[bb.x + i];
            const float4 v = sa.vel[body], w = sa.ang[body];
            lv[i] = v.x; lv[kBodies + i] = v.y; lv[2 * kBodies + i] = v.z;
            lv[3 * kBodies + i] = w.x; lv[4 * kBodies + i] = w.y; lv[5 * kBodies + i] = w.z;
        }
        for (uint32_t i = t; i < nc; i += kBlock) {
            const float4 a = sa.imp[2 * (jb + i)], b = sa.imp[2 * (jb + i) + 1];
            li4[i] = a;
            li2[i] = make_float2(b.x, b.y);
        }
        // Pass cursor: sweep sw, run r, chunk start base, run end end.  Two
        // row buffers alternate (the loop body is written out twice), so a
        // pass's data stays where it was loaded.
        int sw = 0;
        uint32_t r = r0, base = jb, end = runStart[r0 + 1];
        uint32_t j = base + (t >> 1);
        const uint32_t j0 = j < end ? j : jb;
        PairItem bufA = loadLdsRow<kMode>(sa, slotPair[j0], j0, sideB), bufB;
        __syncthreads();
        // Solves the pass in `cur` after starting the next pass's loads into
        // `nxt`; false after the last pass.
        auto pass = [&](PairItem& cur, PairItem& nxt) {
            int nsw = sw;
            uint32_t nr = r, nbase = base + kPass, nend = end;
            const bool runDone = nbase >= end;
            if (runDone) {
                nr = r + 1;
                if (nr == r1) { nr = r0; ++nsw; }
                nbase = runStart[nr];
                nend = runStart[nr + 1];
            }
            const bool more = nsw < sweeps;
            const uint32_t nj = nbase + (t >> 1);
            const uint32_t jl = more && nj < nend ? nj : jb;
            nxt = loadLdsRow<kMode>(sa, slotPair[jl], jl, sideB);
            if (j < end) {
                const uint32_t k = j - jb;
                const float2 e = li2[k];
                cur.i0 = li4[k];
                cur.i1 = make_float4(e.x, e.y, 0.0f, 0.0f);
                if (pairRow<kMode, kBodies, false, false>(sa, j, cur, sideB, lv)) {
                    li4[k] = cur.i0;
                    li2[k] = make_float2(cur.i1.x, cur.i1.y);
                }
            }
            if (runDone) ldsBarrier();
            sw = nsw; r = nr; base = nbase; end = nend; j = nj;
            return more;
        };
        while (pass(bufA, bufB) && pass(bufB, bufA)) {
        }
        for (uint32_t i = t; i < nb; i += kBlock) {
            const uint32_t body = bodyList[bb.x + i];
            sa.vel[body] = make_float4(lv[i], lv[kBodies + i], lv[2 * kBodies + i], 0.0f);
            sa.ang[body] = make_float4(lv[3 * kBodies + i], lv[4 * kBodies + i], lv[5 * kBodies + i], 0.0f);
        }
        if (kMode != 2)
            for (uint32_t i = t; i < nc; i += kBlock) {
                const float2 e = li2[i];
                sa.imp[2 * (jb + i)] = li4[i];
                sa.imp[2 * (jb + i) + 1] = make_float4(e.x, e.y, 0.0f, 0.0f);
            }
        __syncthreads();
    }
    if (t == 0) {
        __threadfence();
        if (atomicAdd(&queue[1], 1u) == gridDim.x - 1) {
            queue[0] = 0u;
            queue[1] = 0u;
        }
    }
}

// Longest-first order of the LDS batches: key = passes per sweep (runs of L
// contacts take ceil(L / kPass) passes), inverted so that an ascending stable
// sort lists the longest first and equal ones in batch order.
__global__ __launch_bounds__(kThreads) void kBatchCostKeys(const uint32_t* __restrict__ runStart,
                                                           const uint32_t* __restrict__ batchFirstRun,
                                                           const uint32_t* __restrict__ ldsBatches, uint32_t kPass,
                                                           uint64_t* __restrict__ keys, uint32_t* __restrict__ vals) {
    const uint32_t q = blockIdx.x * kThreads + threadIdx.x;
    if (q >= *ldsBatches) return;
    uint32_t passes = 0;
    for (uint32_t r = batchFirstRun[q]; r < batchFirstRun[q + 1]; ++r)
        passes += (runStart[r + 1] - runStart[r] + kPass - 1) / kPass;
    keys[q] = 0xffffu - (passes < 0xffffu ? passes : 0xffffu);
    vals[q] = q;
}

// One sweep of the tail colours c0..c1-1 by one workgroup, two lanes per row,
// the rows' written bodies in LDS slots (rowSlot: the tail rows' 32-bit slot
// words, row r0 first).  A colour wider than a pass takes several passes
// with no barrier between them (its rows share no body); each row's data is
// loaded one pass ahead.
template <int kMode>  // 0 two-axis, 1 cone-block
__global__ __launch_bounds__(kSweepRunThreads) void kSweepTailPair(SolveArrays sa, uint32_t base,
                                                                   const uint32_t* __restrict__ colorStart, uint32_t c0,
                                                                   uint32_t c1, const uint32_t* __restrict__ rowSlot,
                                                                   const uint32_t* __restrict__ slotBody,
                                                                   const uint32_t* __restrict__ slotCount) {
    __shared__ float lv[6 * kTailCap];
    __shared__ uint32_t lcs[kTailColors + 1];
    const uint32_t ns = *slotCount;
    if (ns > kTailCap || c1 - c0 > kTailColors) {  // uniform: the global-memory run
        for (uint32_t c = c0; c < c1; ++c) {
            const uint32_t first = colorStart[c], last = colorStart[c + 1];
            for (uint32_t j = first + threadIdx.x; j < last; j += kSweepRunThreads) solveRow<kMode == 1>(sa, base + j);
            __syncthreads();
        }
        return;
    }
    const uint32_t t = threadIdx.x;
    const bool sideB = t & 1;
    constexpr uint32_t kPass = kSweepRunThreads / 2;
    for (uint32_t q = t; q < ns; q += kSweepRunThreads) {
        const uint32_t b = slotBody[q];
        const float4 v = sa.vel[b], w = sa.ang[b];
        lv[q] = v.x; lv[kTailCap + q] = v.y; lv[2 * kTailCap + q] = v.z;
        lv[3 * kTailCap + q] = w.x; lv[4 * kTailCap + q] = w.y; lv[5 * kTailCap + q] = w.z;
    }
    for (uint32_t q = t; q <= c1 - c0; q += kSweepRunThreads) lcs[q] = colorStart[c0 + q];
    const uint32_t f0 = colorStart[c0];  // rowSlot's first row
    // Pass cursors: colour c, rows [cb, ce) of colorStart's numbering; the
    // pass after a colour's last one starts the next colour.  Rows are
    // loaded two passes ahead.
    struct Cursor { uint32_t c, cb, ce; };
    auto next = [&](const Cursor& x) {
        if (x.cb + kPass < x.ce) return Cursor{x.c, x.cb + kPass, x.ce};
        const uint32_t nc = x.c + 1;
        return nc < c1 ? Cursor{nc, lcs[nc - c0], lcs[nc + 1 - c0]} : Cursor{c1, 0u, 0u};
    };
    const uint32_t half = t >> 1;
    Cursor A{c0, f0, colorStart[c0 + 1]};
    PairItem itA, itB;
    if (A.cb + half < A.ce) itA = loadPairItemSlots(sa, rowSlot[A.cb + half - f0], base + A.cb + half, sideB);
    __syncthreads();
    Cursor B = next(A);
    if (B.c < c1 && B.cb + half < B.ce) itB = loadPairItemSlots(sa, rowSlot[B.cb + half - f0], base + B.cb + half, sideB);
    for (;;) {
        const Cursor C = B.c < c1 ? next(B) : Cursor{c1, 0u, 0u};
        PairItem itC;
        if (C.c < c1 && C.cb + half < C.ce)
            itC = loadPairItemSlots(sa, rowSlot[C.cb + half - f0], base + C.cb + half, sideB);
        if (A.cb + half < A.ce) pairRow<kMode, kTailCap>(sa, base + A.cb + half, itA, sideB, lv);
        if (A.cb + kPass >= A.ce) ldsBarrier();
        if (B.c >= c1) break;
        A = B; itA = itB;
        B = C; itB = itC;
    }
    for (uint32_t q = t; q < ns; q += kSweepRunThreads) {
        const uint32_t b = slotBody[q];
        sa.vel[b] = make_float4(lv[q], lv[kTailCap + q], lv[2 * kTailCap + q], 0.0f);
        sa.ang[b] = make_float4(lv[3 * kTailCap + q], lv[4 * kTailCap + q], lv[5 * kTailCap + q], 0.0f);
    }
}

}  // namespace vxk
