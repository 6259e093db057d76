// Device contact rows (src/narrowphase.h:80-366) and velocity integration
// (src/solver.cpp:258-270) for the colored TGS solve.
#pragma once
#include "vxd_narrow.hpp"

namespace vxd {

// solver.cpp:258-270 (scalar / strict path)
VXD void integrateBody(vx_body& b, V3 g, float dt) {
    if (!isDynamic(b) || isLocked(b) || asleepFully(b)) return;
    V3 v = v3(b.velocity);
    v += (g * b.gravityScale + v3(b.force) * solverInvMass(b)) * dt;
    if (!b.fixedRotation) {
        V3 w = v3(b.angularVelocity);
        w += invInertiaMul(b, v3(b.torque)) * dt;
        w *= 1.0f / (1.0f + b.angularDamping * dt);
        b.angularVelocity = raw(w);
    }
    v *= 1.0f / (1.0f + b.linearDamping * dt);
    b.velocity = raw(v);
}

VXD void addVel(vx_body& b, V3 dv) { b.velocity = raw(v3(b.velocity) + dv); }
VXD void subVel(vx_body& b, V3 dv) { b.velocity = raw(v3(b.velocity) - dv); }
VXD void addAng(vx_body& b, V3 dw) { b.angularVelocity = raw(v3(b.angularVelocity) + dw); }
VXD void subAng(vx_body& b, V3 dw) { b.angularVelocity = raw(v3(b.angularVelocity) - dw); }

// narrowphase.h:80-124 (scalar path)
VXD void warmStart(vx_body& a, vx_body& b, vx_contact& c) {
    if (c.normalImpulse <= 0.0f && c.tangentImpulse1 == 0.0f && c.tangentImpulse2 == 0.0f &&
        c.rollingImpulse1 == 0.0f && c.rollingImpulse2 == 0.0f && c.spinningImpulse == 0.0f)
        return;
    V3 pa = anchorWorld(a, v3(c.localAnchorA)), pb = anchorWorld(b, v3(c.localAnchorB));
    c.point = raw((pa + pb) * 0.5f);
    V3 n = v3(c.normal), t1, t2;
    tangents(n, t1, t2);
    V3 imp = n * c.normalImpulse + t1 * c.tangentImpulse1 + t2 * c.tangentImpulse2;
    V3 ang = t1 * c.rollingImpulse1 + t2 * c.rollingImpulse2 + n * c.spinningImpulse;
    if (isDynamic(a)) {
        addVel(a, imp * solverInvMass(a));
        addAng(a, invInertiaMul(a, cross(pa - v3(a.position), imp)));
        addAng(a, invInertiaMul(a, ang));
    }
    if (isDynamic(b)) {
        subVel(b, imp * solverInvMass(b));
        subAng(b, invInertiaMul(b, cross(pb - v3(b.position), imp)));
        subAng(b, invInertiaMul(b, ang));
    }
}

// narrowphase.h:130-366 (scalar path).  Returns the largest impulse change.
VXD float solveContact(vx_body& a, vx_body& b, vx_contact& c, float dt, const vx_solver_options& opt) {
    const float oN = c.normalImpulse, oT1 = c.tangentImpulse1, oT2 = c.tangentImpulse2;
    const float oR1 = c.rollingImpulse1, oR2 = c.rollingImpulse2, oS = c.spinningImpulse;
    V3 pa = anchorWorld(a, v3(c.localAnchorA)), pb = anchorWorld(b, v3(c.localAnchorB));
    c.point = raw((pa + pb) * 0.5f);
    V3 ra = pa - v3(a.position), rb = pb - v3(b.position);
    auto IA = [&](V3 v) { return invInertiaMul(a, v); };
    auto IB = [&](V3 v) { return invInertiaMul(b, v); };
    const V3 n = v3(c.normal);
    V3 raxn = cross(ra, n), rbxn = cross(rb, n);
    float kN = solverInvMass(a) + solverInvMass(b) + dot(raxn, IA(raxn)) + dot(rbxn, IB(rbxn));
    if (kN <= 0.0f) return 0.0f;
    float vn = dot(pointVelocity(a, pa) - pointVelocity(b, pb), n);
    float lg = c.bias0 + dot(n, pa - pb);
    float target = lg > 0.0f ? -lg / dt : 0.0f;
    if (c.vn0 < -1.0f && (lg <= 0.0f || vn > 0.0f)) target = fmx(target, -c.restitution * c.vn0);
    V3 t1, t2;
    tangents(n, t1, t2);
    V3 raxt1 = cross(ra, t1), rbxt1 = cross(rb, t1);
    V3 raxt2 = cross(ra, t2), rbxt2 = cross(rb, t2);
    float k1 = solverInvMass(a) + solverInvMass(b) + dot(raxt1, IA(raxt1)) + dot(rbxt1, IB(rbxt1));
    float k2 = solverInvMass(a) + solverInvMass(b) + dot(raxt2, IA(raxt2)) + dot(rbxt2, IB(rbxt2));
    const float kN1 = dot(raxn, IA(raxt1)) + dot(rbxn, IB(rbxt1));
    const float kN2 = dot(raxn, IA(raxt2)) + dot(rbxn, IB(rbxt2));
    if (opt.frictionModel != VX_FRICTION_CONE_BLOCK) {
        float jn = (target - vn) / kN;
        float nextN = fmx(c.normalImpulse + jn, 0.0f);
        jn = nextN - c.normalImpulse;
        c.normalImpulse = nextN;
        V3 ni = n * jn;
        if (isDynamic(a)) { addVel(a, ni * solverInvMass(a)); addAng(a, IA(cross(ra, ni))); }
        if (isDynamic(b)) { subVel(b, ni * solverInvMass(b)); subAng(b, IB(cross(rb, ni))); }
        if (k1 > 0.0f && k2 > 0.0f) {
            V3 rv = pointVelocity(a, pa) - pointVelocity(b, pb);
            float o1 = c.tangentImpulse1, o2 = c.tangentImpulse2;
            float n1 = o1 - dot(rv, t1) / k1;
            float n2 = o2 - dot(rv, t2) / k2;
            float m1 = c.friction1 * c.normalImpulse, m2 = c.friction2 * c.normalImpulse;
            if (m1 <= 0.0f) n1 = 0.0f;
            if (m2 <= 0.0f) n2 = 0.0f;
            float el = (m1 > 0.0f ? n1 * n1 / (m1 * m1) : 0.0f) + (m2 > 0.0f ? n2 * n2 / (m2 * m2) : 0.0f);
            if (el > 1.0f) { float s = 1.0f / sqrtf(el); n1 *= s; n2 *= s; }
            c.tangentImpulse1 = n1;
            c.tangentImpulse2 = n2;
            V3 fi = t1 * (n1 - o1) + t2 * (n2 - o2);
            if (isDynamic(a)) { addVel(a, fi * solverInvMass(a)); addAng(a, IA(cross(ra, fi))); }
            if (isDynamic(b)) { subVel(b, fi * solverInvMass(b)); subAng(b, IB(cross(rb, fi))); }
        }
    } else {
        float on = c.normalImpulse, o1 = c.tangentImpulse1, o2 = c.tangentImpulse2;
        float nn = on, n1 = o1, n2 = o2;
        if (k1 > 0.0f && k2 > 0.0f) {
            const float k12 = dot(raxt1, IA(raxt2)) + dot(rbxt1, IB(rbxt2));
            const float det = kN * (k1 * k2 - k12 * k12) - kN1 * (kN1 * k2 - kN2 * k12) + kN2 * (kN1 * k12 - kN2 * k1);
            if (fabsf(det) > 1e-12f) {
                const V3 rv = pointVelocity(a, pa) - pointVelocity(b, pb);
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
        const float m1 = c.friction1 * nn, m2 = c.friction2 * nn;
        if (m1 <= 0.0f) n1 = 0.0f;
        if (m2 <= 0.0f) n2 = 0.0f;
        const float el = (m1 > 0.0f ? n1 * n1 / (m1 * m1) : 0.0f) + (m2 > 0.0f ? n2 * n2 / (m2 * m2) : 0.0f);
        if (el > 1.0f) { const float s = 1.0f / sqrtf(el); n1 *= s; n2 *= s; }
        c.normalImpulse = nn;
        c.tangentImpulse1 = n1;
        c.tangentImpulse2 = n2;
        const V3 imp = n * (nn - on) + t1 * (n1 - o1) + t2 * (n2 - o2);
        if (isDynamic(a)) { addVel(a, imp * solverInvMass(a)); addAng(a, IA(cross(ra, imp))); }
        if (isDynamic(b)) { subVel(b, imp * solverInvMass(b)); subAng(b, IB(cross(rb, imp))); }
    }
    V3 wr = v3(a.angularVelocity) - v3(b.angularVelocity);
    float kr1 = dot(t1, IA(t1)) + dot(t1, IB(t1));
    float kr2 = dot(t2, IA(t2)) + dot(t2, IB(t2));
    if (c.rollingFriction > 0.0f && kr1 > 0.0f && kr2 > 0.0f) {
        float o1 = c.rollingImpulse1, o2 = c.rollingImpulse2;
        float n1 = o1 - dot(wr, t1) / kr1;
        float n2 = o2 - dot(wr, t2) / kr2;
        float lim = c.rollingFriction * c.normalImpulse;
        float mag = sqrtf(n1 * n1 + n2 * n2);
        if (mag > lim && mag > 1e-9f) { float s = lim / mag; n1 *= s; n2 *= s; }
        c.rollingImpulse1 = n1;
        c.rollingImpulse2 = n2;
        V3 tq = t1 * (n1 - o1) + t2 * (n2 - o2);
        if (isDynamic(a)) addAng(a, IA(tq));
        if (isDynamic(b)) subAng(b, IB(tq));
    }
    wr = v3(a.angularVelocity) - v3(b.angularVelocity);
    float ks = dot(n, IA(n)) + dot(n, IB(n));
    if (c.spinningFriction > 0.0f && ks > 0.0f) {
        float o = c.spinningImpulse;
        float nx = o - dot(wr, n) / ks;
        float lim = c.spinningFriction * c.normalImpulse;
        nx = fclamp(nx, -lim, lim);
        c.spinningImpulse = nx;
        V3 tq = n * (nx - o);
        if (isDynamic(a)) addAng(a, IA(tq));
        if (isDynamic(b)) subAng(b, IB(tq));
    }
    float mc = fabsf(c.normalImpulse - oN);
    mc = fmx(mc, fabsf(c.tangentImpulse1 - oT1));
    mc = fmx(mc, fabsf(c.tangentImpulse2 - oT2));
    mc = fmx(mc, fabsf(c.rollingImpulse1 - oR1));
    mc = fmx(mc, fabsf(c.rollingImpulse2 - oR2));
    return fmx(mc, fabsf(c.spinningImpulse - oS));
}

}  // namespace vxd
