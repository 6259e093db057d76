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
    float4* rows;     // per contact: 8 float4 (see kPrepareRows)
    uint4* meta;      // per contact: a, b, flags
    float4* imp;      // per contact: 2 float4 (normal, t1, t2, roll1 | roll2, spin)
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

__global__ __launch_bounds__(kThreads) void kLoadImpulses(const vx_contact* __restrict__ cs, uint32_t m, float4* imp) {
    uint32_t k = blockIdx.x * kThreads + threadIdx.x;
    if (k >= m) return;
    const vx_contact& c = cs[k];
    imp[2 * k] = make_float4(c.normalImpulse, c.tangentImpulse1, c.tangentImpulse2, c.rollingImpulse1);
    imp[2 * k + 1] = make_float4(c.rollingImpulse2, c.spinningImpulse, 0.0f, 0.0f);
}

__global__ __launch_bounds__(kThreads) void kStoreImpulses(vx_contact* __restrict__ cs, uint32_t m, const float4* imp) {
    uint32_t k = blockIdx.x * kThreads + threadIdx.x;
    if (k >= m) return;
    float4 a = imp[2 * k], b = imp[2 * k + 1];
    cs[k].normalImpulse = a.x; cs[k].tangentImpulse1 = a.y; cs[k].tangentImpulse2 = a.z; cs[k].rollingImpulse1 = a.w;
    cs[k].rollingImpulse2 = b.x; cs[k].spinningImpulse = b.y;
}

// The pose-dependent prefix of solveContact (narrowphase.h:137-224), once per
// substep for every contact (all colours at once: poses are read-only here).
// Row layout:
//   r0 = ra, kN        r1 = rb, k1          r2 = n, k2         r3 = t1, targetBase
//   r4 = t2, liveGap   r5 = vn0, e, mu1, mu2                   r6 = roll, spin, kr1, kr2
//   r7 = ks, kN1, kN2, k12
__global__ __launch_bounds__(kThreads) void kPrepareRows(const vx_body* __restrict__ bodies, vx_contact* __restrict__ cs,
                                                         uint32_t m, float h, SolveArrays sa) {
    uint32_t k = blockIdx.x * kThreads + threadIdx.x;
    if (k >= m) return;
    const vx_contact c = cs[k];
    const vx_body a = bodies[c.a], b = bodies[c.b];
    uint32_t flags = 0;
    if (isDynamic(a)) flags |= kDynA;
    if (isDynamic(b)) flags |= kDynB;
    if (isDynamic(a) && !rotationLocked(a)) flags |= kRotA;
    if (isDynamic(b) && !rotationLocked(b)) flags |= kRotB;
    float4* row = sa.rows + 8 * (size_t)k;
    if (isSensor(a) || isSensor(b)) {
        sa.meta[k] = make_uint4(c.a, c.b, flags | kSkip | kSensorRow, 0);
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
    sa.meta[k] = make_uint4(c.a, c.b, flags, 0);
}

__device__ inline V3 xyz(float4 v) { return {v.x, v.y, v.z}; }

// Velocity-dependent rows of solveContact (narrowphase.h:183-365) for one
// contact whose pose-dependent rows are already loaded.
struct RowRegs {
    float4 r[8];
    uint32_t flags;
};

template <bool kCone>
__device__ inline void solveRegs(const RowRegs& R, float4 fa, float4 iia, float4 fb, float4 iib, V3& vA, V3& wA, V3& vB,
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
__device__ inline void solveRow(SolveArrays sa, uint32_t k) {
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

template <bool kCone>
__global__ __launch_bounds__(kThreads) void kSweepColor(SolveArrays sa, const uint32_t* __restrict__ colorStart,
                                                        uint32_t color) {
    const uint32_t first = colorStart[color], count = colorStart[color + 1] - first;
    const uint32_t stride = gridDim.x * kThreads;
    for (uint32_t k = blockIdx.x * kThreads + threadIdx.x; k < count; k += stride) solveRow<kCone>(sa, first + k);
}

}  // namespace vxk

namespace vxk {

// Warm start of one contact on the compact arrays (narrowphase.h:80-124);
// anchors come from the substep rows (same arithmetic as recomputing them).
__device__ inline void warmRow(SolveArrays sa, uint32_t k) {
    const uint4 meta = sa.meta[k];
    const uint32_t flags = meta.z;
    const float4 i0 = sa.imp[2 * k], i1 = sa.imp[2 * k + 1];
    const float cN = i0.x, cT1 = i0.y, cT2 = i0.z, cR1 = i0.w, cR2 = i1.x, cS = i1.y;
    if ((flags & kSensorRow) || (cN <= 0.0f && cT1 == 0.0f && cT2 == 0.0f && cR1 == 0.0f && cR2 == 0.0f && cS == 0.0f))
        return;
    const uint32_t ia = meta.x, ib = meta.y;
    const float4* row = sa.rows + 8 * (size_t)k;
    const V3 ra = xyz(row[0]), rb = xyz(row[1]), n = xyz(row[2]), t1 = xyz(row[3]), t2 = xyz(row[4]);
    const bool dynA = flags & kDynA, dynB = flags & kDynB, rotA = flags & kRotA, rotB = flags & kRotB;
    const float4 fa = sa.frame[ia], fb = sa.frame[ib], iia = sa.invI[ia], iib = sa.invI[ib];
    V3 imp = n * cN + t1 * cT1 + t2 * cT2;
    V3 ang = t1 * cR1 + t2 * cR2 + n * cS;
    if (dynA) {
        V3 v = xyz(sa.vel[ia]) + imp * iia.w;
        V3 w = xyz(sa.ang[ia]);
        w += frameMul(fa, iia, cross(ra, imp), rotA);
        w += frameMul(fa, iia, ang, rotA);
        sa.vel[ia] = make_float4(v.x, v.y, v.z, 0.0f);
        sa.ang[ia] = make_float4(w.x, w.y, w.z, 0.0f);
    }
    if (dynB) {
        V3 v = xyz(sa.vel[ib]) - imp * iib.w;
        V3 w = xyz(sa.ang[ib]);
        w -= frameMul(fb, iib, cross(rb, imp), rotB);
        w -= frameMul(fb, iib, ang, rotB);
        sa.vel[ib] = make_float4(v.x, v.y, v.z, 0.0f);
        sa.ang[ib] = make_float4(w.x, w.y, w.z, 0.0f);
    }
}

// One colour of the contacts listed in idx[colorStart[c] .. colorStart[c+1]).
template <int kMode>  // 0 = two-axis sweep, 1 = cone-block sweep, 2 = warm start
__global__ __launch_bounds__(kThreads) void kSweepColorIdx(SolveArrays sa, const uint32_t* __restrict__ idx,
                                                           const uint32_t* __restrict__ colorStart, uint32_t color) {
    const uint32_t first = colorStart[color], count = colorStart[color + 1] - first;
    const uint32_t stride = gridDim.x * kThreads;
    for (uint32_t j = blockIdx.x * kThreads + threadIdx.x; j < count; j += stride) {
        const uint32_t k = idx[first + j];
        if (kMode == 2) warmRow(sa, k);
        else solveRow<kMode == 1>(sa, k);
    }
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
    if (k >= m) return;
    const uint4 mt = meta[k];
    uint32_t root = kNone;
    if (mt.z & kDynA) root = ufFind(parent, mt.x);
    else if (mt.z & kDynB) root = ufFind(parent, mt.y);
    islandOf[k] = root;
    if (root != kNone) atomicAdd(&size[root], 1u);
}

constexpr uint32_t kTinyIsland = 64;

// Splits contacts (in colour order) into small islands (sorted by island,
// colour order kept) and the big-island list (colour order kept).
__global__ __launch_bounds__(kThreads) void kIslandClassify(const uint32_t* __restrict__ islandOf,
                                                            const uint32_t* __restrict__ size, uint32_t m,
                                                            uint32_t limit, uint32_t* __restrict__ bigFlag,
                                                            uint64_t* __restrict__ smallKeys,
                                                            uint32_t* __restrict__ smallVals,
                                                            uint32_t* __restrict__ smallCounter, int mb, int nb) {
    uint32_t k = blockIdx.x * kThreads + threadIdx.x;
    uint32_t root = k < m ? islandOf[k] : kNone;
    const bool big = root != kNone && size[root] > limit;
    const bool small = root != kNone && !big;
    if (k < m) bigFlag[k] = big ? 1u : 0u;
    uint32_t at = waveReserve(small ? 1u : 0u, smallCounter);
    if (small) {
        // Tiny islands (one wavefront's worth of contacts) sort first.
        const uint64_t medium = size[root] > kTinyIsland ? 1u : 0u;
        smallKeys[at] = (medium << (nb + mb)) | ((uint64_t)root << mb) | k;
        smallVals[at] = k;
    }
}

__global__ __launch_bounds__(kThreads) void kBigList(const uint32_t* __restrict__ bigFlag,
                                                     const uint32_t* __restrict__ bigPos, uint32_t m,
                                                     const uint32_t* __restrict__ color, uint32_t* __restrict__ bigIdx,
                                                     uint32_t* __restrict__ bigColorCount) {
    uint32_t k = blockIdx.x * kThreads + threadIdx.x;
    if (k >= m || !bigFlag[k]) return;
    bigIdx[bigPos[k]] = k;
    atomicAdd(&bigColorCount[color[k]], 1u);
}

// Run heads of the island-sorted small list: a run is a maximal stretch of
// one island and one colour.
__global__ __launch_bounds__(kThreads) void kRunHeads(const uint64_t* __restrict__ keys, const uint32_t* __restrict__ idx,
                                                      uint32_t ms, const uint32_t* __restrict__ color,
                                                      uint32_t* __restrict__ runHead, uint32_t* __restrict__ islandHead,
                                                      int mb, int nb, uint32_t* __restrict__ tinyIslands) {
    uint32_t j = blockIdx.x * kThreads + threadIdx.x;
    if (j >= ms) return;
    bool newIsland = j == 0 || (keys[j] >> mb) != (keys[j - 1] >> mb);
    bool newRun = newIsland || color[idx[j]] != color[idx[j - 1]];
    runHead[j] = newRun ? 1u : 0u;
    islandHead[j] = newIsland ? 1u : 0u;
    if (newIsland && (keys[j] >> (nb + mb)) == 0) atomicAdd(tinyIslands, 1u);
}

__global__ __launch_bounds__(kThreads) void kRunTables(const uint32_t* __restrict__ runHead,
                                                       const uint32_t* __restrict__ islandHead,
                                                       const uint32_t* __restrict__ runPos,
                                                       const uint32_t* __restrict__ islandPos, uint32_t ms,
                                                       uint32_t* __restrict__ runStart,
                                                       uint32_t* __restrict__ islandFirstRun) {
    uint32_t j = blockIdx.x * kThreads + threadIdx.x;
    if (j >= ms) return;
    if (runHead[j]) runStart[runPos[j]] = j;
    if (islandHead[j]) islandFirstRun[islandPos[j]] = runPos[j];
}

__global__ void kRunSentinels(uint32_t* runStart, const uint32_t* runs, uint32_t ms, uint32_t* islandFirstRun,
                              const uint32_t* islands, const uint32_t* tiny, uint32_t* ranges) {
    runStart[*runs] = ms;
    islandFirstRun[*islands] = *runs;
    ranges[0] = 0; ranges[1] = *tiny;         // tiny islands: one wavefront each
    ranges[2] = *tiny; ranges[3] = *islands;  // the rest: one workgroup each
}

// Small islands, one workgroup each: every sweep of every colour of the
// island, barrier between colours (the island shares no dynamic body with
// any other contact, so this is the colour order exactly).
// Islands [first, last) of the island list, kBlock threads each (one
// wavefront for tiny islands).
template <int kMode, int kBlock>
__global__ __launch_bounds__(kBlock) void kIslandSweeps(SolveArrays sa, const uint32_t* __restrict__ idx,
                                                        const uint32_t* __restrict__ runStart,
                                                        const uint32_t* __restrict__ islandFirstRun,
                                                        const uint32_t* __restrict__ islandRange, int sweeps) {
    const uint32_t first = islandRange[0], last = islandRange[1];
    for (uint32_t isl = first + blockIdx.x; isl < last; isl += gridDim.x) {
        const uint32_t r0 = islandFirstRun[isl], r1 = islandFirstRun[isl + 1];
        for (int s = 0; s < sweeps; ++s)
            for (uint32_t r = r0; r < r1; ++r) {
                const uint32_t j0 = runStart[r], j1 = runStart[r + 1];
                for (uint32_t j = j0 + threadIdx.x; j < j1; j += kBlock) {
                    const uint32_t k = idx[j];
                    if (kMode == 2) warmRow(sa, k);
                    else solveRow<kMode == 1>(sa, k);
                }
                __syncthreads();
            }
    }
}


// Tiny islands (at most 64 contacts), one wavefront each: every lane owns one
// contact and keeps its rows and impulses in registers for all sweeps; the
// island's bodies live in LDS (each body once, found by scanning the
// wavefront's endpoint list), so a colour hand-off costs an LDS round trip
// instead of global memory.  Same per-body update order as the colour sweep.
template <int kMode>  // 0 two-axis, 1 cone-block, 2 warm start
__global__ __launch_bounds__(64) void kTinyIslands(SolveArrays sa, const uint32_t* __restrict__ idx,
                                                   const uint32_t* __restrict__ runStart,
                                                   const uint32_t* __restrict__ islandFirstRun,
                                                   const uint32_t* __restrict__ islandRange, int sweeps) {
    __shared__ uint32_t bodyId[128];
    __shared__ float4 bodyVel[128], bodyAng[128];
    const uint32_t first = islandRange[0], last = islandRange[1];
    const uint32_t t = threadIdx.x;
    for (uint32_t isl = first + blockIdx.x; isl < last; isl += gridDim.x) {
        const uint32_t r0 = islandFirstRun[isl], r1 = islandFirstRun[isl + 1];
        const uint32_t j0 = runStart[r0], j1 = runStart[r1];
        const bool active = j0 + t < j1;
        uint32_t k = 0, run = 0;
        uint4 meta = make_uint4(kNone, kNone, kSkip, 0);
        if (active) {
            k = idx[j0 + t];
            meta = sa.meta[k];
            for (uint32_t r = r0; r < r1; ++r)
                if (runStart[r] <= j0 + t) run = r - r0;
        }
        bodyId[2 * t] = active && (meta.z & kDynA) ? meta.x : kNone;
        bodyId[2 * t + 1] = active && (meta.z & kDynB) ? meta.y : kNone;
        __syncthreads();
        // Canonical slot of each endpoint: the first slot holding the body.
        uint32_t slotA = 2 * t, slotB = 2 * t + 1;
        if (active) {
            for (uint32_t q = 0; q < 2 * t + 2; ++q) {
                if (q < slotA && bodyId[q] == bodyId[2 * t]) slotA = q;
                if (q < slotB && bodyId[q] == bodyId[2 * t + 1]) slotB = q;
            }
            if (slotA == 2 * t && bodyId[slotA] != kNone) {
                bodyVel[slotA] = sa.vel[meta.x];
                bodyAng[slotA] = sa.ang[meta.x];
            }
            if (slotB == 2 * t + 1 && bodyId[slotB] != kNone) {
                bodyVel[slotB] = sa.vel[meta.y];
                bodyAng[slotB] = sa.ang[meta.y];
            }
        }
        RowRegs R;
        R.flags = meta.z;
        float4 fa = make_float4(0, 0, 0, 1), fb = fa, iia = make_float4(0, 0, 0, 0), iib = iia;
        float4 i0 = make_float4(0, 0, 0, 0), i1 = i0;
        if (active) {
            for (int q = 0; q < 8; ++q) R.r[q] = sa.rows[8 * (size_t)k + q];
            fa = sa.frame[meta.x]; iia = sa.invI[meta.x];
            fb = sa.frame[meta.y]; iib = sa.invI[meta.y];
            i0 = sa.imp[2 * k]; i1 = sa.imp[2 * k + 1];
        }
        float cN = i0.x, cT1 = i0.y, cT2 = i0.z, cR1 = i0.w, cR2 = i1.x, cS = i1.y;
        // Static endpoints keep their (unchanging) global velocities.
        const V3 vStaticA = active && !(meta.z & kDynA) ? xyz(sa.vel[meta.x]) : V3{};
        const V3 wStaticA = active && !(meta.z & kDynA) ? xyz(sa.ang[meta.x]) : V3{};
        const V3 vStaticB = active && !(meta.z & kDynB) ? xyz(sa.vel[meta.y]) : V3{};
        const V3 wStaticB = active && !(meta.z & kDynB) ? xyz(sa.ang[meta.y]) : V3{};
        __syncthreads();
        const uint32_t runs = r1 - r0;
        for (int sw = 0; sw < sweeps; ++sw)
            for (uint32_t r = 0; r < runs; ++r) {
                if (active && run == r) {
                    const bool dynA = meta.z & kDynA, dynB = meta.z & kDynB;
                    V3 vA = dynA ? xyz(bodyVel[slotA]) : vStaticA, wA = dynA ? xyz(bodyAng[slotA]) : wStaticA;
                    V3 vB = dynB ? xyz(bodyVel[slotB]) : vStaticB, wB = dynB ? xyz(bodyAng[slotB]) : wStaticB;
                    if (kMode == 2) {
                        if (!(meta.z & kSensorRow) &&
                            !(cN <= 0.0f && cT1 == 0.0f && cT2 == 0.0f && cR1 == 0.0f && cR2 == 0.0f && cS == 0.0f)) {
                            const V3 ra = xyz(R.r[0]), rb = xyz(R.r[1]), n = xyz(R.r[2]), t1 = xyz(R.r[3]), t2 = xyz(R.r[4]);
                            V3 imp = n * cN + t1 * cT1 + t2 * cT2;
                            V3 ang = t1 * cR1 + t2 * cR2 + n * cS;
                            if (dynA) {
                                vA = vA + imp * iia.w;
                                wA += frameMul(fa, iia, cross(ra, imp), meta.z & kRotA);
                                wA += frameMul(fa, iia, ang, meta.z & kRotA);
                            }
                            if (dynB) {
                                vB = vB - imp * iib.w;
                                wB -= frameMul(fb, iib, cross(rb, imp), meta.z & kRotB);
                                wB -= frameMul(fb, iib, ang, meta.z & kRotB);
                            }
                        }
                    } else if (!(meta.z & kSkip)) {
                        solveRegs<kMode == 1>(R, fa, iia, fb, iib, vA, wA, vB, wB, cN, cT1, cT2, cR1, cR2, cS);
                    }
                    if (dynA) { bodyVel[slotA] = make_float4(vA.x, vA.y, vA.z, 0.0f); bodyAng[slotA] = make_float4(wA.x, wA.y, wA.z, 0.0f); }
                    if (dynB) { bodyVel[slotB] = make_float4(vB.x, vB.y, vB.z, 0.0f); bodyAng[slotB] = make_float4(wB.x, wB.y, wB.z, 0.0f); }
                }
                __syncthreads();
            }
        if (active) {
            if (slotA == 2 * t && bodyId[slotA] != kNone) { sa.vel[meta.x] = bodyVel[slotA]; sa.ang[meta.x] = bodyAng[slotA]; }
            if (slotB == 2 * t + 1 && bodyId[slotB] != kNone) { sa.vel[meta.y] = bodyVel[slotB]; sa.ang[meta.y] = bodyAng[slotB]; }
            sa.imp[2 * k] = make_float4(cN, cT1, cT2, cR1);
            sa.imp[2 * k + 1] = make_float4(cR2, cS, 0.0f, 0.0f);
        }
        __syncthreads();
    }
}

}  // namespace vxk
