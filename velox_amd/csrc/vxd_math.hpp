// Device-side FP32 vector, quaternion and body math for the HIP device step.
//
// Every expression keeps Velox's operation order (include/velox/math.h,
// include/velox/body.h) and the library is compiled with -ffp-contract=off and
// correctly rounded division and square root, so device results are bitwise
// equal to the reference's strict CPU build.
#pragma once
#include <hip/hip_runtime.h>
#ifndef VXD
#define VXD __device__ inline
#endif
#include <cfloat>
#include <cmath>
#include <cstdint>
#include "../../include/vx_records.h"

namespace vxd {

struct V3 { float x = 0, y = 0, z = 0; };
struct Q4 { float x = 0, y = 0, z = 0, w = 1; };

VXD V3 v3(const vx_vec3& v) { return {v.x, v.y, v.z}; }
VXD vx_vec3 raw(const V3& v) { return {v.x, v.y, v.z}; }
VXD Q4 q4(const vx_quat& q) { return {q.x, q.y, q.z, q.w}; }
VXD vx_quat raw(const Q4& q) { return {q.x, q.y, q.z, q.w}; }

// math.h:20-33
VXD V3 operator+(V3 a, V3 b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
VXD V3 operator-(V3 a, V3 b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
VXD V3 operator*(V3 a, float s) { return {a.x * s, a.y * s, a.z * s}; }
VXD V3 operator-(V3 a) { return {-a.x, -a.y, -a.z}; }
VXD V3& operator+=(V3& a, V3 b) { a.x += b.x; a.y += b.y; a.z += b.z; return a; }
VXD V3& operator-=(V3& a, V3 b) { a.x -= b.x; a.y -= b.y; a.z -= b.z; return a; }
VXD V3& operator*=(V3& a, float s) { a.x *= s; a.y *= s; a.z *= s; return a; }

// math.h:36-49: cross() orders its operands so a signed-zero result is canonical.
VXD float dot(V3 a, V3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
// The sign of an exactly-zero component is + when a > b lexicographically
// (a component that differs decides; a NaN compares as not greater) or a == b,
// and - otherwise.  Written with selects only: no branches in the hot loops.
VXD V3 cross(V3 a, V3 b) {
    const bool ex = a.x != b.x, ey = a.y != b.y, ez = a.z != b.z;
    const bool lx = !(a.x > b.x), ly = !(a.y > b.y), lz = !(a.z > b.z);
    const bool neg = (ex & lx) | (!ex & ey & ly) | (!ex & !ey & ez & lz);
    const float zero = neg ? -0.0f : 0.0f;
    const float x = a.y * b.z - a.z * b.y;
    const float y = a.z * b.x - a.x * b.z;
    const float z = a.x * b.y - a.y * b.x;
    return {x != 0.0f ? x : zero, y != 0.0f ? y : zero, z != 0.0f ? z : zero};
}
VXD float lengthSq(V3 v) { return dot(v, v); }
VXD float length(V3 v) { return sqrtf(lengthSq(v)); }
VXD V3 normalize(V3 v) {  // math.h:52-55
    float len = length(v);
    return len > 1e-8f ? v * (1.0f / len) : V3{};
}
VXD float fmn(float a, float b) { return a < b ? a : b; }  // math.h:56-57
VXD float fmx(float a, float b) { return a > b ? a : b; }
VXD float fclamp(float v, float lo, float hi) { return fmn(fmx(v, lo), hi); }
VXD V3 vmin3(V3 a, V3 b) { return {fmn(a.x, b.x), fmn(a.y, b.y), fmn(a.z, b.z)}; }
VXD V3 vmax3(V3 a, V3 b) { return {fmx(a.x, b.x), fmx(a.y, b.y), fmx(a.z, b.z)}; }

// math.h:64-72: unit-quaternion products snap to a 2^-22 grid.
VXD float snapGrid(float value) {
    const float scaled = value * 4194304.0f;
    const float r = scaled >= 0.0f ? floorf(scaled + 0.5f) : ceilf(scaled - 0.5f);
    return r / 4194304.0f;
}

// math.h:74-86
VXD Q4 qmul(Q4 a, Q4 b) {
    Q4 r{a.w * b.x + a.x * b.w + a.y * b.z - a.z * b.y,
         a.w * b.y - a.x * b.z + a.y * b.w + a.z * b.x,
         a.w * b.z + a.x * b.y - a.y * b.x + a.z * b.w,
         a.w * b.w - a.x * b.x - a.y * b.y - a.z * b.z};
    const float na = a.x * a.x + a.y * a.y + a.z * a.z + a.w * a.w;
    const float nb = b.x * b.x + b.y * b.y + b.z * b.z + b.w * b.w;
    const bool unitA = !(na < 0.999f || na > 1.001f);
    const bool unitB = !(nb < 0.999f || nb > 1.001f);
    if (!(unitA && unitB)) return r;
    return {snapGrid(r.x), snapGrid(r.y), snapGrid(r.z), snapGrid(r.w)};
}

// math.h:88-111
VXD Q4 qnormalize(Q4 q) {
    const float n2 = q.x * q.x + q.y * q.y + q.z * q.z + q.w * q.w;
    if (n2 >= 1e-16f && n2 <= FLT_MAX) {
        const float inv = 1.0f / sqrtf(n2);
        return {q.x * inv, q.y * inv, q.z * inv, q.w * inv};
    }
    const float s = fmx(fmx(fabsf(q.x), fabsf(q.y)), fmx(fabsf(q.z), fabsf(q.w)));
    if (!(s >= 1e-8f)) return {0, 0, 0, 0};
    const float x = q.x / s, y = q.y / s, z = q.z / s, w = q.w / s;
    const float len = sqrtf(x * x + y * y + z * z + w * w);
    if (!(len >= 1e-8f)) return {0, 0, 0, 0};
    const float inv = 1.0f / len;
    return {x * inv, y * inv, z * inv, w * inv};
}

// cross() without the canonical zero sign.  Only for an operand of another
// cross(): cross() reads its operands through products, whose zero results it
// canonicalises, and comparisons, which see -0 == +0, so its result does not
// depend on the signs of zero operand components.  A zero component of u x v
// that is added to v * w only gives a zero of either sign where the sum is
// zero, so the sum feeds the outer cross() the same operand up to such signs.
VXD V3 crossRaw(V3 a, V3 b) {
    return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
}

// math.h:113-122 (the inner product through crossRaw: same result bit for bit)
VXD V3 rotate(Q4 q, V3 v) {
    V3 u{q.x, q.y, q.z};
    V3 t = cross(u, crossRaw(u, v) + v * q.w);
    return v + t * 2.0f;
}
VXD V3 rotateInv(Q4 q, V3 v) { return rotate(Q4{-q.x, -q.y, -q.z, q.w}, v); }

// math.h:125-128
VXD Q4 qintegrate(Q4 q, V3 w, float dt) {
    Q4 dq = qmul(Q4{w.x * 0.5f * dt, w.y * 0.5f * dt, w.z * 0.5f * dt, 0.0f}, q);
    return qnormalize(Q4{q.x + dq.x, q.y + dq.y, q.z + dq.z, q.w + dq.w});
}

// ---------------------------------------------------------------------------
// Body predicates and inertia (body.h:177-321) on the raw record.
// ---------------------------------------------------------------------------
VXD bool isStatic(const vx_body& b) { return b.motionType == VX_MOTION_STATIC; }
VXD bool isDynamic(const vx_body& b) { return b.motionType == VX_MOTION_DYNAMIC; }
VXD bool isSensor(const vx_body& b) { return b.sensor != 0; }
VXD bool isLocked(const vx_body& b) { return b.ccdTuning.quality == VX_QUALITY_LOCKED; }
VXD bool rotationLocked(const vx_body& b) { return isLocked(b) || b.fixedRotation != 0; }
VXD bool asleepFully(const vx_body& b) { return b.asleep == 1; }  // sleep.h:67
VXD bool inert(const vx_body& b) { return isStatic(b) || asleepFully(b); }

// body.h:196-202
VXD bool canCollide(const vx_body& a, const vx_body& b) {
    if (a.groupIndex != 0 && a.groupIndex == b.groupIndex) return a.groupIndex > 0;
    return (a.maskBits & b.categoryBits) != 0 && (b.maskBits & a.categoryBits) != 0;
}

// body.h:207-209
VXD float solverInvMass(const vx_body& b) {
    return isDynamic(b) && !isLocked(b) ? b.invMass : 0.0f;
}

// body.h:212-217
VXD Q4 inertiaFrameAt(const vx_body& b, Q4 at) {
    const vx_quat& f = b.inertiaOrientation;
    if (f.x == 0.0f && f.y == 0.0f && f.z == 0.0f && f.w == 1.0f) return at;
    return qmul(at, q4(f));
}

// body.h:220-228
VXD V3 invInertiaMulAt(const vx_body& b, V3 v, Q4 at) {
    if (!isDynamic(b) || rotationLocked(b)) return {};
    Q4 frame = inertiaFrameAt(b, at);
    V3 local = rotateInv(frame, v);
    return rotate(frame, V3{local.x * b.invInertia.x, local.y * b.invInertia.y,
                            local.z * b.invInertia.z});
}
VXD V3 invInertiaMul(const vx_body& b, V3 v) { return invInertiaMulAt(b, v, q4(b.orientation)); }

// body.h:236-247
VXD V3 inertiaMul(const vx_body& b, V3 v) {
    if (rotationLocked(b)) return {};
    Q4 frame = inertiaFrameAt(b, q4(b.orientation));
    V3 local = rotateInv(frame, v);
    V3 w{b.invInertia.x > 0.0f ? local.x / b.invInertia.x : 0.0f,
         b.invInertia.y > 0.0f ? local.y / b.invInertia.y : 0.0f,
         b.invInertia.z > 0.0f ? local.z / b.invInertia.z : 0.0f};
    return rotate(frame, w);
}
VXD V3 angularMomentum(const vx_body& b) { return inertiaMul(b, v3(b.angularVelocity)); }

// body.h:266-300
VXD void advanceOrientation(vx_body& b, float dt) {
    if (rotationLocked(b)) { b.angularVelocity = {0, 0, 0}; return; }
    const V3 ii = v3(b.invInertia);
    const bool aniso = ii.x != ii.y || ii.y != ii.z;
    const float lo = fmn(ii.x, fmn(ii.y, ii.z));
    const float hi = fmx(ii.x, fmx(ii.y, ii.z));
    const bool stable = lo > 0.0f && hi / lo <= 1.0e6f;
    if (isDynamic(b) && aniso && stable) {
        const V3 L = angularMomentum(b);
        const Q4 q0 = q4(b.orientation);
        const V3 w0 = v3(b.angularVelocity);
        Q4 q1 = qintegrate(q0, w0, dt);
        for (int it = 0; it < 3; ++it) {
            V3 w1 = invInertiaMulAt(b, L, q1);
            V3 wm = (w0 + w1) * 0.5f;
            q1 = qintegrate(q0, wm, dt);
        }
        b.orientation = raw(q1);
        b.angularVelocity = raw(invInertiaMul(b, L));
    } else {
        b.orientation = raw(qintegrate(q4(b.orientation), v3(b.angularVelocity), dt));
    }
}

// body.h:306-321
VXD void advanceTransform(vx_body& b, float dt) {
    if (isLocked(b)) return;
    const V3 p0 = v3(b.position);
    const V3 d = v3(b.velocity) * dt;
    V3 p = p0 + d;
    if (d.x != 0.0f && p.x == p0.x) p.x = nextafterf(p0.x, d.x > 0.0f ? FLT_MAX : -FLT_MAX);
    if (d.y != 0.0f && p.y == p0.y) p.y = nextafterf(p0.y, d.y > 0.0f ? FLT_MAX : -FLT_MAX);
    if (d.z != 0.0f && p.z == p0.z) p.z = nextafterf(p0.z, d.z > 0.0f ? FLT_MAX : -FLT_MAX);
    b.position = raw(p);
    advanceOrientation(b, dt);
}

// body.h:330-344
VXD float maxPointSpeed(const vx_body& b) {
    float r = b.radius;
    const V3 he = v3(b.halfExtents);
    switch (b.shape) {
    case VX_SHAPE_BOX: r = length(he); break;
    case VX_SHAPE_ROUNDED_BOX: r = length(he) + b.radius; break;
    case VX_SHAPE_ELLIPSOID: r = fmx(he.x, fmx(he.y, he.z)); break;
    case VX_SHAPE_CAPSULE: r = b.capsuleHalfHeight + b.radius; break;
    case VX_SHAPE_CYLINDER:
        r = sqrtf(b.capsuleHalfHeight * b.capsuleHalfHeight + b.radius * b.radius); break;
    case VX_SHAPE_CONE: {
        float top = 1.5f * b.capsuleHalfHeight;
        r = fmx(top, sqrtf(b.radius * b.radius + 0.25f * b.capsuleHalfHeight * b.capsuleHalfHeight));
        break;
    }
    default: break;
    }
    return length(v3(b.velocity)) + length(v3(b.angularVelocity)) * r;
}

VXD bool sweeps(const vx_body& b) {  // narrowphase.h:386-388, 434-439
    return b.ccdTuning.enableContinuous && b.ccdTuning.quality != VX_QUALITY_LOW &&
           b.ccdTuning.quality != VX_QUALITY_LOCKED;
}

}  // namespace vxd
