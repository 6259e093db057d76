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

enum RowFlag : uint32_t { kDynA = 1, kDynB = 2, kRotA = 4, kRotB = 8, kSkip = 16 };

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
        sa.meta[k] = make_uint4(c.a, c.b, flags | kSkip, 0);
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

// One colour of one sweep: the velocity-dependent rows of solveContact
// (narrowphase.h:183-365) for contacts [colorStart[color], colorStart[color+1]).
template <bool kCone>
__device__ inline void solveRow(SolveArrays sa, uint32_t k) {
    const uint4 meta = sa.meta[k];
    const uint32_t flags = meta.z;
    if (flags & kSkip) return;
    const uint32_t ia = meta.x, ib = meta.y;
    const float4* row = sa.rows + 8 * (size_t)k;
    const float4 r0 = row[0], r1 = row[1], r2 = row[2], r3 = row[3], r4 = row[4], r5 = row[5], r6 = row[6], r7 = row[7];
    const V3 ra = xyz(r0), rb = xyz(r1), n = xyz(r2), t1 = xyz(r3), t2 = xyz(r4);
    const float kN = r0.w, k1 = r1.w, k2 = r2.w, targetBase = r3.w, lg = r4.w;
    const float vn0 = r5.x, rest = r5.y, mu1 = r5.z, mu2 = r5.w;
    const float rolling = r6.x, spinning = r6.y, kr1 = r6.z, kr2 = r6.w;
    const float ks = r7.x, kN1 = r7.y, kN2 = r7.z, k12 = r7.w;
    const bool dynA = flags & kDynA, dynB = flags & kDynB, rotA = flags & kRotA, rotB = flags & kRotB;
    const float4 fa = sa.frame[ia], fb = sa.frame[ib], iia = sa.invI[ia], iib = sa.invI[ib];
    const float simA = iia.w, simB = iib.w;
    auto IA = [&](V3 v) { return frameMul(fa, iia, v, rotA); };
    auto IB = [&](V3 v) { return frameMul(fb, iib, v, rotB); };
    V3 vA = xyz(sa.vel[ia]), wA = xyz(sa.ang[ia]), vB = xyz(sa.vel[ib]), wB = xyz(sa.ang[ib]);
    float4 i0 = sa.imp[2 * k], i1 = sa.imp[2 * k + 1];
    float cN = i0.x, cT1 = i0.y, cT2 = i0.z, cR1 = i0.w, cR2 = i1.x, cS = i1.y;

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
            if (el > 1.0f) { float s = 1.0f / sqrtf(el); n1 *= s; n2 *= s; }
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
        if (el > 1.0f) { const float s = 1.0f / sqrtf(el); n1 *= s; n2 *= s; }
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
        if (mag > lim && mag > 1e-9f) { float s = lim / mag; n1 *= s; n2 *= s; }
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
    if (dynA) { sa.vel[ia] = make_float4(vA.x, vA.y, vA.z, 0.0f); sa.ang[ia] = make_float4(wA.x, wA.y, wA.z, 0.0f); }
    if (dynB) { sa.vel[ib] = make_float4(vB.x, vB.y, vB.z, 0.0f); sa.ang[ib] = make_float4(wB.x, wB.y, wB.z, 0.0f); }
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
