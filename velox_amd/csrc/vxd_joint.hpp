// Device restatement of the reference's joint solver (src/joint_solver.h),
// the per-joint routine its CUDA backend runs colour by colour
// (backend_cuda.cu:329-339).  Same operations in the same order as the
// reference, so results agree bit for bit where the reference uses only
// arithmetic and sqrtf; atan2f and acosf (hinge, cone-twist and wheel
// steering limits, six-degree-of-freedom and soft-weld rotation vectors) are
// the device's libm.
#pragma once
#include "vxd_math.hpp"

namespace vxd {

constexpr float kJointBeta = 0.2f;            // joint_solver.h:7
constexpr float kJointMaxPositionBias = 10.0f;  // joint_solver.h:8
constexpr int kJointIterations = 8;           // joint_solver.h:6

VXD V3 jointClampBias(V3 bias) {  // joint_solver.h:10-14
    const float m = lengthSq(bias);
    if (m <= kJointMaxPositionBias * kJointMaxPositionBias) return bias;
    return bias * (kJointMaxPositionBias / sqrtf(m));
}

struct M3 { V3 c0, c1, c2; };

VXD V3 mul(const M3& m, V3 v) { return m.c0 * v.x + m.c1 * v.y + m.c2 * v.z; }

VXD M3 inverse(const M3& m) {  // joint_solver.h:24-33
    const V3 r0 = cross(m.c1, m.c2), r1 = cross(m.c2, m.c0), r2 = cross(m.c0, m.c1);
    const float det = dot(m.c0, r0);
    const float inv = fabsf(det) > 1e-12f ? 1.0f / det : 0.0f;
    return {{r0.x * inv, r1.x * inv, r2.x * inv}, {r0.y * inv, r1.y * inv, r2.y * inv},
            {r0.z * inv, r1.z * inv, r2.z * inv}};
}

VXD M3 pointMass(const vx_body& a, const vx_body& b, V3 ra, V3 rb) {  // joint_solver.h:35-48
    const V3 axes[3] = {{1, 0, 0}, {0, 1, 0}, {0, 0, 1}};
    V3 col[3];
    for (int i = 0; i < 3; ++i) {
        const V3 e = axes[i];
        V3 k = e * (solverInvMass(a) + solverInvMass(b));
        k = k + cross(invInertiaMul(a, cross(ra, e)), ra);
        k = k + cross(invInertiaMul(b, cross(rb, e)), rb);
        col[i] = k;
    }
    return {col[0], col[1], col[2]};
}

VXD M3 angularMass(const vx_body& a, const vx_body& b) {  // joint_solver.h:50-54
    return {invInertiaMul(a, {1, 0, 0}) + invInertiaMul(b, {1, 0, 0}),
            invInertiaMul(a, {0, 1, 0}) + invInertiaMul(b, {0, 1, 0}),
            invInertiaMul(a, {0, 0, 1}) + invInertiaMul(b, {0, 0, 1})};
}

VXD V3 anchorVelocity(const vx_body& b, V3 r) { return v3(b.velocity) + cross(v3(b.angularVelocity), r); }

VXD void jointImpulse(vx_body& body, V3 r, V3 impulse, float sign) {  // joint_solver.h:60-65
    if (!isDynamic(body) || isLocked(body)) return;
    body.velocity = raw(v3(body.velocity) + impulse * (sign * solverInvMass(body)));
    body.angularVelocity = raw(v3(body.angularVelocity) + invInertiaMul(body, cross(r, impulse)) * sign);
}

VXD void addTo(vx_vec3& acc, V3 v) { acc = raw(v3(acc) + v); }

VXD void applyLinear(vx_joint& j, vx_body& a, vx_body& b, V3 ra, V3 rb, V3 impulse, float signA) {
    jointImpulse(a, ra, impulse, signA);
    jointImpulse(b, rb, impulse, -signA);
    addTo(j.reactionLinearImpulse, impulse * signA);
}

VXD void applyAngular(vx_joint& j, vx_body& a, vx_body& b, V3 impulse, float signA = 1.0f) {
    if (isDynamic(a) && !isLocked(a)) a.angularVelocity = raw(v3(a.angularVelocity) + invInertiaMul(a, impulse) * signA);
    if (isDynamic(b) && !isLocked(b)) b.angularVelocity = raw(v3(b.angularVelocity) - invInertiaMul(b, impulse) * signA);
    addTo(j.reactionAngularImpulse, impulse * signA);
}

VXD void jointFrame(const vx_joint& j, const vx_body& a, const vx_body& b, V3& xA, V3& yA, V3& zA, V3& xB, V3& yB,
                    V3& zB) {  // joint_solver.h:82-93
    xA = normalize(rotate(q4(a.orientation), v3(j.localAxisA)));
    yA = normalize(rotate(q4(a.orientation), v3(j.localRefA)));
    zA = normalize(cross(xA, yA));
    yA = cross(zA, xA);
    xB = normalize(rotate(q4(b.orientation), v3(j.localAxisB)));
    yB = normalize(rotate(q4(b.orientation), v3(j.localRefB)));
    zB = normalize(cross(xB, yB));
    yB = cross(zB, xB);
}

VXD Q4 matrixQuaternion(float m00, float m01, float m02, float m10, float m11, float m12, float m20, float m21,
                        float m22) {  // joint_solver.h:95-117
    Q4 q;
    const float trace = m00 + m11 + m22;
    if (trace > 0.0f) {
        const float s = sqrtf(trace + 1.0f) * 2.0f;
        q = {(m21 - m12) / s, (m02 - m20) / s, (m10 - m01) / s, 0.25f * s};
    } else if (m00 > m11 && m00 > m22) {
        const float s = sqrtf(1.0f + m00 - m11 - m22) * 2.0f;
        q = {0.25f * s, (m01 + m10) / s, (m02 + m20) / s, (m21 - m12) / s};
    } else if (m11 > m22) {
        const float s = sqrtf(1.0f + m11 - m00 - m22) * 2.0f;
        q = {(m01 + m10) / s, 0.25f * s, (m12 + m21) / s, (m02 - m20) / s};
    } else {
        const float s = sqrtf(1.0f + m22 - m00 - m11) * 2.0f;
        q = {(m02 + m20) / s, (m12 + m21) / s, 0.25f * s, (m10 - m01) / s};
    }
    return qnormalize(q);
}

VXD V3 rotationVector(const vx_joint& j, const vx_body& a, const vx_body& b) {  // joint_solver.h:119-134
    V3 xA, yA, zA, xB, yB, zB;
    jointFrame(j, a, b, xA, yA, zA, xB, yB, zB);
    Q4 q = matrixQuaternion(dot(xA, xB), dot(xA, yB), dot(xA, zB), dot(yA, xB), dot(yA, yB), dot(yA, zB), dot(zA, xB),
                            dot(zA, yB), dot(zA, zB));
    if (q.w < 0.0f) q = {-q.x, -q.y, -q.z, -q.w};
    const V3 v{q.x, q.y, q.z};
    const float sinHalf = length(v);
    if (sinHalf < 1e-7f) return v * 2.0f;
    const float angle = 2.0f * atan2f(sinHalf, fclamp(q.w, 0.0f, 1.0f));
    return v * (angle / sinHalf);
}

VXD float comp(V3 v, int axis) { return axis == 0 ? v.x : (axis == 1 ? v.y : v.z); }
VXD float& comp(vx_vec3& v, int axis) { return axis == 0 ? v.x : (axis == 1 ? v.y : v.z); }

VXD void jointReset(vx_joint& j) {  // joint_solver.h:144-157
    j.motorImpulse = 0.0f; j.limitImpulse = 0.0f; j.swingImpulse = 0.0f; j.twistImpulse = 0.0f; j.springImpulse = 0.0f;
    j.linearMotorImpulse = {0, 0, 0}; j.angularMotorImpulse = {0, 0, 0};
    j.linearLimitImpulse = {0, 0, 0}; j.angularLimitImpulse = {0, 0, 0};
    j.reactionLinearImpulse = {0, 0, 0}; j.reactionAngularImpulse = {0, 0, 0};
    j.broken = 0;
}

VXD void solveAngularFrame(vx_joint& j, vx_body& a, vx_body& b, float biasScale) {  // joint_solver.h:159-168
    V3 xA, yA, zA, xB, yB, zB;
    jointFrame(j, a, b, xA, yA, zA, xB, yB, zB);
    const V3 error = (cross(xB, xA) + cross(yB, yA) + cross(zB, zA)) * 0.5f;
    const V3 relW = v3(a.angularVelocity) - v3(b.angularVelocity);
    const V3 impulse = mul(inverse(angularMass(a, b)), -(relW + error * biasScale));
    applyAngular(j, a, b, impulse);
}

VXD void solveSixDof(vx_joint& j, vx_body& a, vx_body& b, V3 ra, V3 rb, V3 pa, V3 pb, float dt) {  // 170-245
    V3 xA, yA, zA, xB, yB, zB;
    jointFrame(j, a, b, xA, yA, zA, xB, yB, zB);
    const V3 axes[3] = {xA, yA, zA};
    const V3 d = pb - pa;
    const M3 pointK = pointMass(a, b, ra, rb);
    for (int i = 0; i < 3; ++i) {
        const uint8_t bit = (uint8_t)(1u << i);
        const V3 axis = axes[i];
        const float k = dot(axis, mul(pointK, axis));
        if (k <= 0.0f) continue;
        float speed = dot(anchorVelocity(b, rb) - anchorVelocity(a, ra), axis);
        if (j.linearMotorMask & bit) {
            float lambda = (comp(v3(j.linearMotorSpeed), i) - speed) / k;
            const float maxImpulse = comp(v3(j.maxLinearMotorForce), i) * dt;
            float& sum = comp(j.linearMotorImpulse, i);
            const float old = sum;
            sum = fclamp(old + lambda, -maxImpulse, maxImpulse);
            lambda = sum - old;
            applyLinear(j, a, b, ra, rb, axis * lambda, -1.0f);
            speed = dot(anchorVelocity(b, rb) - anchorVelocity(a, ra), axis);
        }
        if (!(j.linearLimitMask & bit)) continue;
        const float position = dot(d, axis);
        const float lower = comp(v3(j.lowerLinearLimit), i), upper = comp(v3(j.upperLinearLimit), i);
        float lambda = 0.0f;
        float& sum = comp(j.linearLimitImpulse, i);
        const float old = sum;
        if (lower == upper) {
            lambda = -(speed + (position - lower) * (kJointBeta / dt)) / k;
            sum += lambda;
        } else if (position < lower) {
            lambda = -(speed + (position - lower) * (kJointBeta / dt)) / k;
            sum = fmx(0.0f, old + lambda);
            lambda = sum - old;
        } else if (position > upper) {
            lambda = -(speed + (position - upper) * (kJointBeta / dt)) / k;
            sum = fmn(0.0f, old + lambda);
            lambda = sum - old;
        }
        if (lambda != 0.0f) applyLinear(j, a, b, ra, rb, axis * lambda, -1.0f);
    }
    const V3 rotation = rotationVector(j, a, b);
    for (int i = 0; i < 3; ++i) {
        const uint8_t bit = (uint8_t)(1u << i);
        const V3 axis = axes[i];
        const float k = dot(axis, invInertiaMul(a, axis)) + dot(axis, invInertiaMul(b, axis));
        if (k <= 0.0f) continue;
        float speed = dot(v3(b.angularVelocity) - v3(a.angularVelocity), axis);
        if (j.angularMotorMask & bit) {
            float lambda = (speed - comp(v3(j.angularMotorSpeed), i)) / k;
            const float maxImpulse = comp(v3(j.maxAngularMotorTorque), i) * dt;
            float& sum = comp(j.angularMotorImpulse, i);
            const float old = sum;
            sum = fclamp(old + lambda, -maxImpulse, maxImpulse);
            lambda = sum - old;
            applyAngular(j, a, b, axis * lambda);
            speed = dot(v3(b.angularVelocity) - v3(a.angularVelocity), axis);
        }
        if (!(j.angularLimitMask & bit)) continue;
        const float angle = comp(rotation, i);
        const float lower = comp(v3(j.lowerAngularLimit), i), upper = comp(v3(j.upperAngularLimit), i);
        float lambda = 0.0f;
        float& sum = comp(j.angularLimitImpulse, i);
        const float old = sum;
        if (lower == upper) {
            lambda = (speed + (angle - lower) * (kJointBeta / dt)) / k;
            sum += lambda;
        } else if (angle < lower) {
            lambda = (speed + (angle - lower) * (kJointBeta / dt)) / k;
            sum = fmn(0.0f, old + lambda);
            lambda = sum - old;
        } else if (angle > upper) {
            lambda = (speed + (angle - upper) * (kJointBeta / dt)) / k;
            sum = fmx(0.0f, old + lambda);
            lambda = sum - old;
        }
        if (lambda != 0.0f) applyAngular(j, a, b, axis * lambda);
    }
}

VXD void solvePrismatic(vx_joint& j, vx_body& a, vx_body& b, V3 ra, V3 rb, V3 pa, V3 pb, float dt) {  // 247-301
    const V3 axis = normalize(rotate(q4(a.orientation), v3(j.localAxisA)));
    V3 ref = rotate(q4(a.orientation), v3(j.localRefA));
    V3 t1 = normalize(ref - axis * dot(ref, axis));
    if (lengthSq(t1) < 1e-12f) {
        ref = fabsf(axis.x) < 0.9f ? V3{1, 0, 0} : V3{0, 1, 0};
        t1 = normalize(cross(axis, ref));
    }
    const V3 tangents[2] = {t1, cross(axis, t1)};
    const M3 pointK = pointMass(a, b, ra, rb);
    const V3 error = pa - pb;
    V3 velocity = anchorVelocity(a, ra) - anchorVelocity(b, rb);
    for (int i = 0; i < 2; ++i) {
        const V3 tangent = tangents[i];
        const float k = dot(tangent, mul(pointK, tangent));
        if (k <= 0.0f) continue;
        const float lambda = -(dot(velocity, tangent) + dot(error, tangent) * (kJointBeta / dt)) / k;
        applyLinear(j, a, b, ra, rb, tangent * lambda, 1.0f);
        velocity = anchorVelocity(a, ra) - anchorVelocity(b, rb);
    }
    solveAngularFrame(j, a, b, kJointBeta / dt);
    const float kAxis = dot(axis, mul(pointK, axis));
    if (kAxis <= 0.0f) return;
    float axisSpeed = dot(anchorVelocity(b, rb) - anchorVelocity(a, ra), axis);
    if (j.enableMotor) {
        float lambda = (j.motorSpeed - axisSpeed) / kAxis;
        const float maxImpulse = j.maxMotorForce * dt;
        const float old = j.motorImpulse;
        j.motorImpulse = fclamp(old + lambda, -maxImpulse, maxImpulse);
        lambda = j.motorImpulse - old;
        applyLinear(j, a, b, ra, rb, axis * lambda, -1.0f);
        axisSpeed = dot(anchorVelocity(b, rb) - anchorVelocity(a, ra), axis);
    }
    if (!j.enableLimit) return;
    const float translation = dot(pb - pa, axis);
    float lambda = 0.0f;
    const float old = j.limitImpulse;
    if (translation < j.lowerLimit) {
        const float bias = (translation - j.lowerLimit) * (kJointBeta / dt);
        lambda = -(axisSpeed + bias) / kAxis;
        j.limitImpulse = fmx(0.0f, old + lambda);
        lambda = j.limitImpulse - old;
    } else if (translation > j.upperLimit) {
        const float bias = (translation - j.upperLimit) * (kJointBeta / dt);
        lambda = -(axisSpeed + bias) / kAxis;
        j.limitImpulse = fmn(0.0f, old + lambda);
        lambda = j.limitImpulse - old;
    }
    if (lambda != 0.0f) applyLinear(j, a, b, ra, rb, axis * lambda, -1.0f);
}

VXD void solveHinge(vx_joint& j, vx_body& a, vx_body& b, V3 ra, V3 rb, V3 pa, V3 pb, float dt) {  // 303-367
    const V3 velocity = anchorVelocity(a, ra) - anchorVelocity(b, rb);
    const V3 bias = jointClampBias((pa - pb) * (kJointBeta / dt));
    const V3 impulse = mul(inverse(pointMass(a, b, ra, rb)), -(velocity + bias));
    applyLinear(j, a, b, ra, rb, impulse, 1.0f);
    const V3 axisA = rotate(q4(a.orientation), v3(j.localAxisA));
    const V3 axisB = rotate(q4(b.orientation), v3(j.localAxisB));
    const V3 error = cross(axisB, axisA);
    const V3 ref = fabsf(axisA.x) < 0.9f ? V3{1, 0, 0} : V3{0, 1, 0};
    V3 tangents[2] = {normalize(cross(axisA, ref)), {0, 0, 0}};
    tangents[1] = cross(axisA, tangents[0]);
    V3 relW = v3(a.angularVelocity) - v3(b.angularVelocity);
    for (int i = 0; i < 2; ++i) {
        const V3 tangent = tangents[i];
        const float k = dot(tangent, invInertiaMul(a, tangent)) + dot(tangent, invInertiaMul(b, tangent));
        if (k <= 0.0f) continue;
        const float lambda = -(dot(relW, tangent) + dot(error, tangent) * (kJointBeta / dt)) / k;
        applyAngular(j, a, b, tangent * lambda);
        relW = v3(a.angularVelocity) - v3(b.angularVelocity);
    }
    const float kAxis = dot(axisA, invInertiaMul(a, axisA)) + dot(axisA, invInertiaMul(b, axisA));
    if (kAxis <= 0.0f) return;
    if (j.enableMotor) {
        const float wAxis = dot(v3(a.angularVelocity) - v3(b.angularVelocity), axisA);
        float lambda = (j.motorSpeed - wAxis) / kAxis;
        const float maxImpulse = j.maxMotorTorque * dt;
        const float old = j.motorImpulse;
        j.motorImpulse = fclamp(old + lambda, -maxImpulse, maxImpulse);
        lambda = j.motorImpulse - old;
        applyAngular(j, a, b, axisA * lambda);
    }
    if (!j.enableLimit) return;
    const V3 refA = rotate(q4(a.orientation), v3(j.localRefA));
    const V3 refB = rotate(q4(b.orientation), v3(j.localRefB));
    const float angle = atan2f(dot(cross(refA, refB), axisA), dot(refA, refB));
    const float wAxis = dot(v3(a.angularVelocity) - v3(b.angularVelocity), axisA);
    float lambda = 0.0f;
    const float old = j.limitImpulse;
    if (angle < j.lowerLimit) {
        const float target = (angle - j.lowerLimit) * (kJointBeta / dt);
        lambda = (target - wAxis) / kAxis;
        j.limitImpulse = fmn(0.0f, old + lambda);
        lambda = j.limitImpulse - old;
    } else if (angle > j.upperLimit) {
        const float target = (angle - j.upperLimit) * (kJointBeta / dt);
        lambda = (target - wAxis) / kAxis;
        j.limitImpulse = fmx(0.0f, old + lambda);
        lambda = j.limitImpulse - old;
    }
    if (lambda != 0.0f) applyAngular(j, a, b, axisA * lambda);
}

VXD void solveConeTwist(vx_joint& j, vx_body& a, vx_body& b, V3 ra, V3 rb, V3 pa, V3 pb, float dt) {  // 369-435
    const V3 velocity = anchorVelocity(a, ra) - anchorVelocity(b, rb);
    const V3 bias = (pa - pb) * (kJointBeta / dt);
    const V3 impulse = mul(inverse(pointMass(a, b, ra, rb)), -(velocity + bias));
    applyLinear(j, a, b, ra, rb, impulse, 1.0f);
    const V3 axisA = normalize(rotate(q4(a.orientation), v3(j.localAxisA)));
    const V3 axisB = normalize(rotate(q4(b.orientation), v3(j.localAxisB)));
    V3 relW = v3(a.angularVelocity) - v3(b.angularVelocity);
    if (j.enableSwingLimit) {
        const float angle = acosf(fclamp(dot(axisA, axisB), -1.0f, 1.0f));
        if (angle > j.swingLimit) {
            V3 n = cross(axisB, axisA);
            if (lengthSq(n) < 1e-10f) {
                const V3 refA = rotate(q4(a.orientation), v3(j.localRefA));
                n = cross(axisA, refA);
            }
            n = normalize(n);
            const float k = dot(n, invInertiaMul(a, n)) + dot(n, invInertiaMul(b, n));
            if (k > 0.0f) {
                const float error = angle - j.swingLimit;
                float lambda = -(dot(relW, n) + error * (kJointBeta / dt)) / k;
                const float old = j.swingImpulse;
                j.swingImpulse = fmn(0.0f, old + lambda);
                lambda = j.swingImpulse - old;
                applyAngular(j, a, b, n * lambda);
                relW = v3(a.angularVelocity) - v3(b.angularVelocity);
            }
        }
    }
    if (!j.enableTwistLimit) return;
    V3 refA = rotate(q4(a.orientation), v3(j.localRefA));
    V3 refB = rotate(q4(b.orientation), v3(j.localRefB));
    refA = normalize(refA - axisA * dot(refA, axisA));
    refB = normalize(refB - axisA * dot(refB, axisA));
    if (lengthSq(refA) < 1e-10f || lengthSq(refB) < 1e-10f) return;
    const float angle = atan2f(dot(cross(refA, refB), axisA), dot(refA, refB));
    const float k = dot(axisA, invInertiaMul(a, axisA)) + dot(axisA, invInertiaMul(b, axisA));
    if (k <= 0.0f) return;
    const float wAxis = dot(v3(a.angularVelocity) - v3(b.angularVelocity), axisA);
    float lambda = 0.0f;
    const float old = j.twistImpulse;
    if (angle < j.lowerTwistLimit) {
        const float target = (angle - j.lowerTwistLimit) * (kJointBeta / dt);
        lambda = (target - wAxis) / k;
        j.twistImpulse = fmn(0.0f, old + lambda);
        lambda = j.twistImpulse - old;
    } else if (angle > j.upperTwistLimit) {
        const float target = (angle - j.upperTwistLimit) * (kJointBeta / dt);
        lambda = (target - wAxis) / k;
        j.twistImpulse = fmx(0.0f, old + lambda);
        lambda = j.twistImpulse - old;
    }
    if (lambda != 0.0f) applyAngular(j, a, b, axisA * lambda);
}

VXD void solveWeld(vx_joint& j, vx_body& a, vx_body& b, V3 ra, V3 rb, V3 pa, V3 pb, float dt) {  // 437-475
    const V3 velocity = anchorVelocity(a, ra) - anchorVelocity(b, rb);
    const V3 bias = jointClampBias((pa - pb) * (kJointBeta / dt));
    const V3 impulse = mul(inverse(pointMass(a, b, ra, rb)), -(velocity + bias));
    applyLinear(j, a, b, ra, rb, impulse, 1.0f);
    if (j.weldFrequencyHz > 0.0f) {
        const V3 rotation = rotationVector(j, a, b);
        const V3 relW = v3(a.angularVelocity) - v3(b.angularVelocity);
        constexpr float tau = 6.28318530718f;
        const float omega = tau * j.weldFrequencyHz;
        const M3 angK = angularMass(a, b);
        const float mass = 1.0f / (dot({1, 0, 0}, invInertiaMul(a, {1, 0, 0})) +
                                   dot({1, 0, 0}, invInertiaMul(b, {1, 0, 0})) + 1e-12f);
        const float stiffness = mass * omega * omega;
        const float damping = 2.0f * mass * j.weldDampingRatio * omega;
        const float denom = fmx(dt * (damping + dt * stiffness), 1e-12f);
        const float softness = 1.0f / denom;
        const V3 angularBias = rotation * (dt * stiffness * softness);
        const V3 lambda = mul(inverse(angK), -(relW + angularBias + v3(j.angularMotorImpulse) * softness));
        addTo(j.angularMotorImpulse, lambda);
        applyAngular(j, a, b, lambda);
    } else {
        solveAngularFrame(j, a, b, kJointBeta / dt);
    }
}

VXD void solveWheel(vx_joint& j, vx_body& a, vx_body& b, V3 ra, V3 rb, V3 pa, V3 pb, float dt) {  // 477-582
    const V3 axisA = normalize(rotate(q4(a.orientation), v3(j.localAxisA)));
    V3 ref = rotate(q4(a.orientation), v3(j.localRefA));
    V3 t1 = normalize(ref - axisA * dot(ref, axisA));
    if (lengthSq(t1) < 1e-12f) {
        ref = fabsf(axisA.x) < 0.9f ? V3{1, 0, 0} : V3{0, 1, 0};
        t1 = normalize(cross(axisA, ref));
    }
    const V3 tangents[2] = {t1, cross(axisA, t1)};
    const M3 pointK = pointMass(a, b, ra, rb);
    const V3 error = pa - pb;
    V3 vel = anchorVelocity(a, ra) - anchorVelocity(b, rb);
    for (int i = 0; i < 2; ++i) {
        const V3 tangent = tangents[i];
        const float k = dot(tangent, mul(pointK, tangent));
        if (k <= 0.0f) continue;
        const float lambda = -(dot(vel, tangent) + dot(error, tangent) * (kJointBeta / dt)) / k;
        applyLinear(j, a, b, ra, rb, tangent * lambda, 1.0f);
        vel = anchorVelocity(a, ra) - anchorVelocity(b, rb);
    }
    const float kAxis = dot(axisA, mul(pointK, axisA));
    if (kAxis > 0.0f) {
        const float axisSpeed = dot(anchorVelocity(b, rb) - anchorVelocity(a, ra), axisA);
        const float displacement = dot(pb - pa, axisA);
        constexpr float tau = 6.28318530718f;
        const float mass = 1.0f / kAxis;
        const float omega = tau * j.suspensionFrequencyHz;
        const float stiffness = mass * omega * omega;
        const float damping = 2.0f * mass * j.suspensionDampingRatio * omega;
        const float softness =
            (damping + dt * stiffness) > 1e-10f ? 1.0f / (dt * (damping + dt * stiffness)) : 0.0f;
        const float bias = displacement * dt * stiffness * softness;
        const float lambda = -(axisSpeed + bias + softness * j.springImpulse) / (kAxis + softness);
        j.springImpulse += lambda;
        applyLinear(j, a, b, ra, rb, axisA * lambda, -1.0f);
    }
    V3 relW = v3(a.angularVelocity) - v3(b.angularVelocity);
    for (int i = 0; i < 2; ++i) {
        const V3 tangent = tangents[i];
        const float k = dot(tangent, invInertiaMul(a, tangent)) + dot(tangent, invInertiaMul(b, tangent));
        if (k <= 0.0f) continue;
        const V3 refA = rotate(q4(a.orientation), v3(j.localRefA));
        const V3 refB = rotate(q4(b.orientation), v3(j.localRefB));
        const V3 frameError = cross(refB, refA);
        const float lambda = -(dot(relW, tangent) + dot(frameError, tangent) * (kJointBeta / dt)) / k;
        applyAngular(j, a, b, tangent * lambda);
        relW = v3(a.angularVelocity) - v3(b.angularVelocity);
    }
    if (j.enableSteering) {
        const V3 steerAxis = normalize(rotate(q4(a.orientation), v3(j.localRefA)));
        const float kSteer = dot(steerAxis, invInertiaMul(a, steerAxis)) + dot(steerAxis, invInertiaMul(b, steerAxis));
        if (kSteer > 0.0f) {
            const V3 refA = rotate(q4(a.orientation), v3(j.localRefA));
            const V3 refB = rotate(q4(b.orientation), v3(j.localRefB));
            const V3 axle = axisA;
            const V3 pA = normalize(refA - axle * dot(refA, axle));
            const V3 pB = normalize(refB - axle * dot(refB, axle));
            const float currentAngle = atan2f(dot(cross(pA, pB), steerAxis), dot(pA, pB));
            const float angleError = j.steeringAngle - currentAngle;
            const float wSteer = dot(v3(b.angularVelocity) - v3(a.angularVelocity), steerAxis);
            const float lambda = (wSteer - angleError * (kJointBeta / dt)) / kSteer;
            applyAngular(j, a, b, steerAxis * lambda);
        }
    }
    if (j.enableMotor) {
        const float kAxle = dot(axisA, invInertiaMul(a, axisA)) + dot(axisA, invInertiaMul(b, axisA));
        if (kAxle > 0.0f) {
            const float wAxle = dot(v3(b.angularVelocity) - v3(a.angularVelocity), axisA);
            float lambda = (wAxle - j.motorSpeed) / kAxle;
            const float maxImpulse = j.maxMotorTorque * dt;
            const float old = j.motorImpulse;
            j.motorImpulse = fclamp(old + lambda, -maxImpulse, maxImpulse);
            lambda = j.motorImpulse - old;
            applyAngular(j, a, b, axisA * lambda);
        }
    }
}

VXD void solveRope(vx_joint& j, vx_body& a, vx_body& b, V3 ra, V3 rb, V3 pa, V3 pb, float dt) {  // 584-607
    const V3 d = pa - pb;
    const float len = length(d);
    if (len < 1e-8f || len <= j.maxLength) return;
    const V3 n = d * (1.0f / len);
    const V3 velocity = anchorVelocity(a, ra) - anchorVelocity(b, rb);
    const V3 raxn = cross(ra, n), rbxn = cross(rb, n);
    const float k = solverInvMass(a) + solverInvMass(b) + dot(raxn, invInertiaMul(a, raxn)) +
                    dot(rbxn, invInertiaMul(b, rbxn));
    if (k <= 0.0f) return;
    const float overshoot = len - j.maxLength;
    const float bias = overshoot * (kJointBeta / dt);
    float lambda = -(dot(velocity, n) + bias) / k;
    const float old = j.limitImpulse;
    j.limitImpulse = fmn(0.0f, old + lambda);
    lambda = j.limitImpulse - old;
    applyLinear(j, a, b, ra, rb, n * lambda, 1.0f);
}

VXD void solvePulley(vx_joint& j, vx_body& a, vx_body& b, V3 ra, V3 rb, V3 pa, V3 pb, float dt) {  // 609-651
    const V3 dA = pa - v3(j.pulleyAnchorA), dB = pb - v3(j.pulleyAnchorB);
    const float lenA = length(dA), lenB = length(dB);
    if (lenA < 1e-8f || lenB < 1e-8f) return;
    const V3 nA = dA * (1.0f / lenA), nB = dB * (1.0f / lenB);
    const V3 raxnA = cross(ra, nA), rbxnB = cross(rb, nB);
    const float kA = solverInvMass(a) + dot(raxnA, invInertiaMul(a, raxnA));
    const float kB = solverInvMass(b) + dot(rbxnB, invInertiaMul(b, rbxnB));
    const float ratio = j.pulleyRatio;
    const float k = kA + ratio * ratio * kB;
    if (k <= 0.0f) return;
    const float restTotal = j.restLength;
    const float C = lenA + ratio * lenB - restTotal;
    const float bias = C * (kJointBeta / dt);
    const float speedA = dot(anchorVelocity(a, ra), nA);
    const float speedB = dot(anchorVelocity(b, rb), nB);
    const float cdot = speedA + ratio * speedB;
    float lambda = -(cdot + bias) / k;
    const float old = j.limitImpulse;
    j.limitImpulse += lambda;
    lambda = j.limitImpulse - old;
    jointImpulse(a, ra, nA * lambda, 1.0f);
    jointImpulse(b, rb, nB * (lambda * ratio), 1.0f);
    addTo(j.reactionLinearImpulse, nA * lambda);
}

VXD void solveGear(vx_joint& j, vx_body& a, vx_body& b, float dt) {  // 653-682
    const V3 axisA = normalize(rotate(q4(a.orientation), v3(j.localAxisA)));
    const V3 axisB = normalize(rotate(q4(b.orientation), v3(j.localAxisB)));
    const float ratio = j.gearRatio;
    const float kA = dot(axisA, invInertiaMul(a, axisA));
    const float kB = dot(axisB, invInertiaMul(b, axisB));
    const float k = kA + ratio * ratio * kB;
    if (k <= 0.0f) return;
    const float wA = dot(v3(a.angularVelocity), axisA);
    const float wB = dot(v3(b.angularVelocity), axisB);
    const float cdot = wA + ratio * wB;
    float lambda = -cdot / k;
    const float old = j.motorImpulse;
    j.motorImpulse += lambda;
    lambda = j.motorImpulse - old;
    if (isDynamic(a) && !isLocked(a)) a.angularVelocity = raw(v3(a.angularVelocity) + invInertiaMul(a, axisA * lambda));
    if (isDynamic(b) && !isLocked(b))
        b.angularVelocity = raw(v3(b.angularVelocity) + invInertiaMul(b, axisB * (lambda * ratio)));
    addTo(j.reactionAngularImpulse, axisA * lambda);
    (void)dt;
}

VXD void solveJoint(vx_joint& j, vx_body& a, vx_body& b, float dt) {  // joint_solver.h:684-775
    const V3 ra = rotate(q4(a.orientation), v3(j.localAnchorA));
    const V3 rb = rotate(q4(b.orientation), v3(j.localAnchorB));
    const V3 pa = v3(a.position) + ra, pb = v3(b.position) + rb;
    const V3 velocity = anchorVelocity(a, ra) - anchorVelocity(b, rb);
    switch (j.type) {
    case VX_JOINT_BALL: {
        const V3 bias = (pa - pb) * (kJointBeta / dt);
        const V3 impulse = mul(inverse(pointMass(a, b, ra, rb)), -(velocity + bias));
        applyLinear(j, a, b, ra, rb, impulse, 1.0f);
        break;
    }
    case VX_JOINT_DISTANCE: {
        const V3 d = pa - pb;
        const float len = length(d);
        if (len < 1e-8f) break;
        const V3 n = d * (1.0f / len);
        const V3 raxn = cross(ra, n), rbxn = cross(rb, n);
        const float k = solverInvMass(a) + solverInvMass(b) + dot(raxn, invInertiaMul(a, raxn)) +
                        dot(rbxn, invInertiaMul(b, rbxn));
        if (k <= 0.0f) break;
        float lambda;
        if (j.enableSpring) {
            constexpr float tau = 6.28318530718f;
            const float mass = 1.0f / k;
            const float omega = tau * j.springFrequencyHz;
            const float stiffness = mass * omega * omega;
            const float damping = 2.0f * mass * j.springDampingRatio * omega;
            const float softness = 1.0f / (dt * (damping + dt * stiffness));
            const float bias = (len - j.restLength) * dt * stiffness * softness;
            lambda = -(dot(velocity, n) + bias + softness * j.springImpulse) / (k + softness);
            j.springImpulse += lambda;
        } else {
            const float bias = (len - j.restLength) * (kJointBeta / dt);
            lambda = -(dot(velocity, n) + bias) / k;
        }
        applyLinear(j, a, b, ra, rb, n * lambda, 1.0f);
        break;
    }
    case VX_JOINT_FIXED: {
        const V3 bias = (pa - pb) * (kJointBeta / dt);
        const V3 impulse = mul(inverse(pointMass(a, b, ra, rb)), -(velocity + bias));
        applyLinear(j, a, b, ra, rb, impulse, 1.0f);
        solveAngularFrame(j, a, b, kJointBeta / dt);
        break;
    }
    case VX_JOINT_MOTOR: {
        const V3 bias = (pa - pb) * (kJointBeta / dt);
        V3 impulse = mul(inverse(pointMass(a, b, ra, rb)), -(velocity + bias));
        const float maxImpulse = j.maxMotorForce * dt;
        const float mag = length(impulse);
        if (mag > maxImpulse && mag > 1e-9f) impulse = impulse * (maxImpulse / mag);
        applyLinear(j, a, b, ra, rb, impulse, 1.0f);
        solveAngularFrame(j, a, b, kJointBeta / dt);
        break;
    }
    case VX_JOINT_WELD: solveWeld(j, a, b, ra, rb, pa, pb, dt); break;
    case VX_JOINT_WHEEL: solveWheel(j, a, b, ra, rb, pa, pb, dt); break;
    case VX_JOINT_ROPE: solveRope(j, a, b, ra, rb, pa, pb, dt); break;
    case VX_JOINT_PULLEY: solvePulley(j, a, b, ra, rb, pa, pb, dt); break;
    case VX_JOINT_GEAR: solveGear(j, a, b, dt); break;
    case VX_JOINT_SIX_DOF: solveSixDof(j, a, b, ra, rb, pa, pb, dt); break;
    case VX_JOINT_PRISMATIC: solvePrismatic(j, a, b, ra, rb, pa, pb, dt); break;
    case VX_JOINT_HINGE: solveHinge(j, a, b, ra, rb, pa, pb, dt); break;
    case VX_JOINT_CONE_TWIST: solveConeTwist(j, a, b, ra, rb, pa, pb, dt); break;
    default: break;
    }
}

}  // namespace vxd
