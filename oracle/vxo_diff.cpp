// Double-precision forward and reverse of the differentiable form's frozen-
// active-set step, for the checks of the device adjoint (test infrastructure:
// the product never loads this).
//
// One world at a time.  Its contacts per step come from a recorded device
// rollout (vxhip_world_record): solve order, geometry (normal, local anchors,
// detection gap, friction coefficients, restitution, approach speed) and the
// previous step's contact whose final impulses each one starts from.  The
// forward restates the device step on that data in double precision:
// velocity integration with gravity (CpuBackend::integrate, solver.cpp:224-272,
// no force or damping), per substep the pose-dependent row prefix, the warm
// start (first substep) and 2I sweeps of solveContact's two-axis rows in
// solve order (narrowphase.h:80-233), the transform advance of isotropic
// bodies (body.h:306-321), then the positional correction
// (world.cpp:4364-4397).  The row, warm-start, transform and correction maps
// are vxd_adjoint.hpp's templates at double precision; their adjoints are
// pinned by central differences there (tests/adjoint_check.cpp), and this
// file's forward by central differences of the whole rollout.
//
// Parameters per dynamic body: initial linear and angular velocity, friction
// (entering each contact's coefficients as mu_pair s with mu_pair the
// geometric mean), mass (inverse mass 1/m, inverse inertia k/m).
#include <array>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <vector>
#include "../include/vx_records.h"
#include "../velox_amd/csrc/vxd_adjoint.hpp"

namespace {
using D = double;
using namespace vxadj;

struct Body {
    V<D> x, v, w;
    Q<D> q;
    D im, ii, mu;  // inverse mass, isotropic inverse inertia, friction
    bool dyn, sphere;
};

struct Ct {  // one recorded contact
    uint32_t a, b;  // world-local body indices
    V<D> n, t1, t2, lA, lB;
    D bias0, s1, s2, rest, vn0;
    int32_t match;  // index into the previous step's contacts, or -1
};

void tangents(V<D> n, V<D>& t1, V<D>& t2) {  // narrowphase.h:38-50 (normal frozen: tangents are data too)
    const D ax = std::fabs(n.x), ay = std::fabs(n.y), az = std::fabs(n.z);
    const V<D> ref = ax < ay ? (ax < az ? V<D>{1, 0, 0} : V<D>{0, 0, 1}) : (ay < az ? V<D>{0, 1, 0} : V<D>{0, 0, 1});
    V<D> c = cross(n, ref);
    const D l = std::sqrt(dot(c, c));
    t1 = l > 1e-8 ? c * (1.0 / l) : V<D>{0, 0, 0};
    t2 = cross(n, t1);
}

Q<D> qnorm(Q<D> q) {
    const D l = std::sqrt(q.x * q.x + q.y * q.y + q.z * q.z + q.w * q.w);
    return {q.x / l, q.y / l, q.z / l, q.w / l};
}

V<D> arm(const Body& b, V<D> l) { return b.sphere ? l : rotate(b.q, l); }

struct Problem {
    std::vector<Body> init;            // bodies at the rollout's start
    std::vector<std::vector<Ct>> steps;
    V<D> g;
    D dt;
    int S, sweeps, passes;
    V<D> target;
};

// Per substep tape of a step for the reverse.
struct StepTape {
    std::vector<Body> start;                          // before the step
    std::vector<std::vector<Body>> pre;               // per substep: before the solve
    std::vector<std::vector<Body>> preT;              // per substep: before the transform
    std::vector<Body> postSub;                        // before the correction
    std::vector<std::vector<Visit<D>>> visits;        // per substep: (sweeps + 1) * m visits, warm last
    std::vector<std::vector<std::array<D, 2>>> corr;  // per pass, per contact: g, fired
    std::vector<std::array<D, 3>> lam0;               // warm-start impulses
};

Row<D> makeRow(const Ct& c, const Body& A, const Body& B, D h) {
    Row<D> r;
    r.ra = arm(A, c.lA);
    r.rb = arm(B, c.lB);
    r.n = c.n; r.t1 = c.t1; r.t2 = c.t2;
    r.lg = c.bias0 + dot(c.n, (A.x + r.ra) - (B.x + r.rb));
    auto k = [&](V<D> d) {
        const V<D> ca = cross(r.ra, d), cb = cross(r.rb, d);
        return (A.dyn ? A.im : 0.0) + (B.dyn ? B.im : 0.0) + (A.dyn ? A.ii * dot(ca, ca) : 0.0) +
               (B.dyn ? B.ii * dot(cb, cb) : 0.0);
    };
    r.kN = k(c.n); r.k1 = k(c.t1); r.k2 = k(c.t2);
    r.h = h; r.vn0 = c.vn0; r.rest = c.rest;
    const D mu = std::sqrt(A.mu) * std::sqrt(B.mu);
    r.mu1 = mu * c.s1; r.mu2 = mu * c.s2;
    return r;
}
Side<D> side(const Body& b) { return {b.dyn ? b.im : 0.0, b.dyn ? b.ii : 0.0, b.dyn}; }

// One step forward; returns the final impulses; fills the tape when given.
std::vector<std::array<D, 3>> stepForward(const Problem& P, std::vector<Body>& bodies, const std::vector<Ct>& cs,
                                          const std::vector<std::array<D, 3>>& prevLam, StepTape* tape) {
    const D h = P.dt / P.S;
    const size_t m = cs.size();
    if (tape) tape->start = bodies;
    for (Body& b : bodies) {
        b.q = qnorm(b.q);
        if (b.dyn) b.v += P.g * h;
    }
    std::vector<std::array<D, 3>> lam(m);
    for (size_t k = 0; k < m; ++k)
        lam[k] = cs[k].match >= 0 ? prevLam[cs[k].match] : std::array<D, 3>{0, 0, 0};
    if (tape) {
        tape->lam0 = lam;
        tape->pre.assign(P.S, {});
        tape->preT.assign(P.S, {});
        tape->visits.assign(P.S, std::vector<Visit<D>>((P.sweeps + 1) * m));
    }
    for (int st = 0; st < P.S; ++st) {
        if (st > 0)
            for (Body& b : bodies)
                if (b.dyn) b.v += P.g * h;
        if (tape) tape->pre[st] = bodies;
        std::vector<Row<D>> rows(m);
        for (size_t k = 0; k < m; ++k) rows[k] = makeRow(cs[k], bodies[cs[k].a], bodies[cs[k].b], h);
        auto visitOf = [&](size_t k) {
            const Body &A = bodies[cs[k].a], &B = bodies[cs[k].b];
            return Visit<D>{A.v, A.w, B.v, B.w, lam[k][0], lam[k][1], lam[k][2]};
        };
        auto store = [&](size_t k, const Visit<D>& s) {
            Body &A = bodies[cs[k].a], &B = bodies[cs[k].b];
            if (A.dyn) { A.v = s.vA; A.w = s.wA; }
            if (B.dyn) { B.v = s.vB; B.w = s.wB; }
            lam[k] = {s.cN, s.c1, s.c2};
        };
        if (st == 0)
            for (size_t k = 0; k < m; ++k) {
                Visit<D> s = visitOf(k);
                if (tape) tape->visits[st][P.sweeps * m + k] = s;
                warmForward(rows[k], side(bodies[cs[k].a]), side(bodies[cs[k].b]), s);
                store(k, s);
            }
        for (int it = 0; it < P.sweeps; ++it)
            for (size_t k = 0; k < m; ++k) {
                Visit<D> s = visitOf(k);
                if (tape) tape->visits[st][it * m + k] = s;
                if (!(rows[k].kN > 0)) continue;
                rowForward(rows[k], side(bodies[cs[k].a]), side(bodies[cs[k].b]), s);
                store(k, s);
            }
        if (tape) tape->preT[st] = bodies;
        for (Body& b : bodies)
            if (b.dyn) {
                b.x += b.v * h;
                b.q = quatStep(b.q, b.w, h);
            }
    }
    if (tape) {
        tape->postSub = bodies;
        tape->corr.assign(P.passes, std::vector<std::array<D, 2>>(m));
    }
    for (int it = 0; it < P.passes; ++it)
        for (size_t k = 0; k < m; ++k) {
            const Ct& c = cs[k];
            Body &A = bodies[c.a], &B = bodies[c.b];
            const D imA = A.dyn ? A.im : 0.0, imB = B.dyn ? B.im : 0.0;
            const D g = c.bias0 + dot(c.n, (A.x + arm(A, c.lA)) - (B.x + arm(B, c.lB)));
            const bool fired = correctionFires(g, imA, imB);
            if (tape) tape->corr[it][k] = {g, fired ? 1.0 : 0.0};
            if (!fired) continue;
            const D push = correctionPush(g, imA, imB);
            if (A.dyn) A.x += c.n * (push * imA);
            if (B.dyn) B.x -= c.n * (push * imB);
        }
    return lam;
}

struct Grad {
    std::vector<V<D>> x, v, w;
    std::vector<Q<D>> q;
    std::vector<D> im, ii, mu, mass;
};

D lossOf(const Problem& P, const std::vector<Body>& b, V<D>* comOut) {
    D M = 0;
    V<D> c{0, 0, 0};
    for (const Body& x : b)
        if (x.dyn) { M += 1.0 / x.im; c += x.x * (1.0 / x.im); }
    c = c * (1.0 / M);
    if (comOut) *comOut = c;
    const V<D> d = c - P.target;
    return dot(d, d);
}

// Reverse of one step given the adjoints of its end state and final impulses.
void stepReverse(const Problem& P, const std::vector<Ct>& cs, const StepTape& T, Grad& G,
                 std::vector<std::array<D, 3>>& lamBar) {
    const D h = P.dt / P.S;
    const size_t m = cs.size();
    // positional correction
    for (int it = P.passes - 1; it >= 0; --it)
        for (size_t k = m; k-- > 0;) {
            if (T.corr[it][k][1] == 0.0) continue;
            const Ct& c = cs[k];
            const Body &A = T.postSub[c.a], &B = T.postSub[c.b];
            D imAb = 0, imBb = 0;
            const D gb = correctionAdj(T.corr[it][k][0], A.dyn ? A.im : 0.0, B.dyn ? B.im : 0.0, A.dyn, B.dyn, c.n,
                                       G.x[c.a], G.x[c.b], imAb, imBb);
            if (A.dyn) {
                G.x[c.a] += c.n * gb;
                if (!A.sphere) rotateAdjQ(A.q, c.lA, c.n * gb, G.q[c.a]);
                G.im[c.a] += imAb;
            }
            if (B.dyn) {
                G.x[c.b] -= c.n * gb;
                if (!B.sphere) rotateAdjQ(B.q, c.lB, c.n * (-gb), G.q[c.b]);
                G.im[c.b] += imBb;
            }
        }
    for (int st = P.S - 1; st >= 0; --st) {
        // transform
        const std::vector<Body>& pt = T.preT[st];
        for (size_t i = 0; i < pt.size(); ++i) {
            if (!pt[i].dyn) continue;
            V<D> xb{0, 0, 0};
            Q<D> qb{0, 0, 0, 0};
            transformAdj(pt[i].q, pt[i].w, h, G.x[i], G.q[i], xb, qb, G.v[i], G.w[i]);
            G.x[i] = xb;
            G.q[i] = qb;
        }
        // sweeps, then warm start
        const std::vector<Body>& pb = T.pre[st];
        std::vector<Row<D>> rows(m);
        std::vector<RowBar<D>> rbar(m);
        for (size_t k = 0; k < m; ++k) {
            rows[k] = makeRow(cs[k], pb[cs[k].a], pb[cs[k].b], h);
            clear(rbar[k]);
        }
        auto visit = [&](size_t k, int pass, bool warm) {
            const Ct& c = cs[k];
            const Body &A = pb[c.a], &B = pb[c.b];
            if (!warm && !(rows[k].kN > 0)) return;
            const Visit<D>& pre = T.visits[st][pass * m + k];
            Visit<D> bar{G.v[c.a], G.w[c.a], G.v[c.b], G.w[c.b], lamBar[k][0], lamBar[k][1], lamBar[k][2]};
            if (warm) warmAdjoint(rows[k], side(A), side(B), pre, bar, rbar[k]);
            else rowAdjoint(rows[k], side(A), side(B), pre, bar, rbar[k]);
            if (A.dyn) { G.v[c.a] = bar.vA; G.w[c.a] = bar.wA; }
            if (B.dyn) { G.v[c.b] = bar.vB; G.w[c.b] = bar.wB; }
            lamBar[k] = {bar.cN, bar.c1, bar.c2};
        };
        for (int it = P.sweeps - 1; it >= 0; --it)
            for (size_t k = m; k-- > 0;) visit(k, it, false);
        if (st == 0)
            for (size_t k = m; k-- > 0;) visit(k, P.sweeps, true);
        // row prefix
        for (size_t k = 0; k < m; ++k) {
            const Ct& c = cs[k];
            const Body &A = pb[c.a], &B = pb[c.b];
            RowBar<D> rb = rbar[k];
            const D fricBar = rb.mu1 * rows[k].mu1 + rb.mu2 * rows[k].mu2;
            V<D> xAb{0, 0, 0}, xBb{0, 0, 0};
            rowPrefixAdj(rows[k], side(A), side(B), rb, xAb, xBb);
            if (A.dyn) {
                G.x[c.a] += xAb;
                if (!A.sphere) rotateAdjQ(A.q, c.lA, rb.ra, G.q[c.a]);
                G.im[c.a] += rb.imA;
                G.ii[c.a] += rb.iiA;
                G.mu[c.a] += fricBar / (2.0 * A.mu);
            }
            if (B.dyn) {
                G.x[c.b] += xBb;
                if (!B.sphere) rotateAdjQ(B.q, c.lB, rb.rb, G.q[c.b]);
                G.im[c.b] += rb.imB;
                G.ii[c.b] += rb.iiB;
                G.mu[c.b] += fricBar / (2.0 * B.mu);
            }
        }
    }
    for (size_t i = 0; i < T.start.size(); ++i)
        if (T.start[i].dyn) normalizeAdj(T.start[i].q, G.q[i]);
}

}  // namespace

extern "C" {

// Rollout of one world on recorded contacts.  bodies: the world's n records
// at the start (parameters already applied); contacts: all steps' records,
// step t at [offsets[t], offsets[t+1]), body indices world-local; match:
// per contact, the previous step's world-local contact index or -1.
// Returns the loss; with grad (8 doubles per body: v0.xyz, w0.xyz, friction,
// mass) also its gradient.
// params (optional, 8 doubles per body as in grad) replace the records'
// initial velocities, friction and mass (inverse inertia scaled with it).
double vxo_diff_rollout(const vx_body* bodies, uint32_t n, const vx_contact* contacts, const int32_t* match,
                        const uint32_t* offsets, uint32_t steps, float dt, int substeps, int iterations,
                        int positionIterations, const float gravity[3], const double target[3], const double* params,
                        double* grad, double* comOut) {
    Problem P;
    P.g = {gravity[0], gravity[1], gravity[2]};
    P.dt = dt;
    P.S = substeps;
    P.sweeps = 2 * iterations;
    P.passes = positionIterations;
    P.target = {target[0], target[1], target[2]};
    P.init.resize(n);
    for (uint32_t i = 0; i < n; ++i) {
        const vx_body& r = bodies[i];
        Body& b = P.init[i];
        b.x = {r.position.x, r.position.y, r.position.z};
        b.q = {r.orientation.x, r.orientation.y, r.orientation.z, r.orientation.w};
        b.v = {r.velocity.x, r.velocity.y, r.velocity.z};
        b.w = {r.angularVelocity.x, r.angularVelocity.y, r.angularVelocity.z};
        b.dyn = r.motionType == 2;
        b.im = r.invMass;
        b.ii = r.invInertia.x;
        b.mu = r.friction;
        b.sphere = r.shape == VX_SHAPE_SPHERE;
        if (params && b.dyn) {
            const double* p = params + 8 * (size_t)i;
            b.v = {p[0], p[1], p[2]};
            b.w = {p[3], p[4], p[5]};
            b.mu = p[6];
            const D kappa = (D)r.invInertia.x / (D)r.invMass;
            b.im = 1.0 / p[7];
            b.ii = kappa / p[7];
        }
    }
    P.steps.resize(steps);
    for (uint32_t t = 0; t < steps; ++t)
        for (uint32_t k = offsets[t]; k < offsets[t + 1]; ++k) {
            const vx_contact& r = contacts[k];
            Ct c;
            c.a = r.a; c.b = r.b;
            c.n = {r.normal.x, r.normal.y, r.normal.z};
            tangents(c.n, c.t1, c.t2);
            c.lA = {r.localAnchorA.x, r.localAnchorA.y, r.localAnchorA.z};
            c.lB = {r.localAnchorB.x, r.localAnchorB.y, r.localAnchorB.z};
            c.bias0 = r.bias0;
            const D mu = std::sqrt((D)bodies[r.a].friction) * std::sqrt((D)bodies[r.b].friction);
            c.s1 = mu > 0 ? r.friction1 / mu : 0.0;
            c.s2 = mu > 0 ? r.friction2 / mu : 0.0;
            c.rest = r.restitution;
            c.vn0 = r.vn0;
            c.match = match[k];
            P.steps[t].push_back(c);
        }
    std::vector<Body> b = P.init;
    std::vector<StepTape> tapes(grad ? steps : 0);
    std::vector<std::array<D, 3>> lam;
    for (uint32_t t = 0; t < steps; ++t) lam = stepForward(P, b, P.steps[t], lam, grad ? &tapes[t] : nullptr);
    V<D> com;
    const D L = lossOf(P, b, &com);
    if (comOut) { comOut[0] = com.x; comOut[1] = com.y; comOut[2] = com.z; }
    if (!grad) return L;
    Grad G;
    G.x.assign(n, {0, 0, 0}); G.v.assign(n, {0, 0, 0}); G.w.assign(n, {0, 0, 0});
    G.q.assign(n, {0, 0, 0, 0});
    G.im.assign(n, 0); G.ii.assign(n, 0); G.mu.assign(n, 0); G.mass.assign(n, 0);
    D M = 0;
    for (const Body& x : b) if (x.dyn) M += 1.0 / x.im;
    const V<D> d = com - P.target;
    for (uint32_t i = 0; i < n; ++i)
        if (b[i].dyn) {
            const D mass = 1.0 / b[i].im;
            G.x[i] = d * (2.0 * mass / M);
            G.mass[i] += 2.0 / M * dot(d, b[i].x - com);
        }
    std::vector<std::array<D, 3>> lamBar(P.steps.empty() ? 0 : P.steps.back().size(), {0, 0, 0});
    for (uint32_t t = steps; t-- > 0;) {
        stepReverse(P, P.steps[t], tapes[t], G, lamBar);
        std::vector<std::array<D, 3>> prev(t ? P.steps[t - 1].size() : 0, {0, 0, 0});
        for (size_t k = 0; k < P.steps[t].size(); ++k)
            if (P.steps[t][k].match >= 0)
                for (int q = 0; q < 3; ++q) prev[P.steps[t][k].match][q] += lamBar[k][q];
        lamBar.swap(prev);
    }
    for (uint32_t i = 0; i < n; ++i) {
        double* o = grad + 8 * (size_t)i;
        if (!P.init[i].dyn) { for (int q = 0; q < 8; ++q) o[q] = 0; continue; }
        o[0] = G.v[i].x; o[1] = G.v[i].y; o[2] = G.v[i].z;
        o[3] = G.w[i].x; o[4] = G.w[i].y; o[5] = G.w[i].z;
        o[6] = G.mu[i];
        const D im = P.init[i].im, ii = P.init[i].ii;
        o[7] = G.mass[i] - G.im[i] * im * im - G.ii[i] * ii * im;
    }
    return L;
}

}  // extern "C"
