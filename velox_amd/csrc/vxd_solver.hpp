// Device velocity integration (src/solver.cpp:258-270).  The contact rows of
// the colored TGS solve (src/narrowphase.h:80-366) are in vx_solve.hpp, split
// into a pose-dependent prefix per substep and the per-sweep velocity part.
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

}  // namespace vxd
