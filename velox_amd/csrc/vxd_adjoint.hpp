// Reverse-mode (adjoint) derivatives of the device step's substepped contact
// solve and integration, with the active set frozen per step.
//
// The forward maps differentiated here are the step's own, written for the
// bodies the differentiable form admits (dynamic bodies with isotropic
// inverse inertia: the inverse inertia tensor is invI * identity in every
// orientation, so I^-1 v = invI v; static bodies take no update):
//
//  * one sequential-impulse row of solveContact's two-axis path
//    (narrowphase.h:130-233): normal row with the live-gap target and the
//    restitution floor, accumulated clamp at zero, then both friction rows
//    with the elliptic Coulomb clamp, each applied to both bodies;
//  * warmStartContact (narrowphase.h:80-124);
//  * the pose-dependent row prefix (narrowphase.h:137-224): arms
//    r = R(q) l, live gap bias0 + n.(pa - pb), effective masses
//    m_a + m_b + invI_a |r_a x d|^2 + invI_b |r_b x d|^2;
//  * the transform advance (body.h:306-321 with the isotropic branch of
//    advanceOrientation: q' = normalize(q + (w h / 2, 0) q)), whose
//    forward-progress guard and canonical rounding are taken as identity;
//  * positional correction (world.cpp:4364-4397).
//
// "Frozen active set": each contact's bodies, normal, tangents, local
// anchors, detection gap (bias0), restitution and approach speed (vn0) are
// the step's data, not functions of the state; the branches the forward took
// (normal clamp, friction zeroing and ellipse scaling, restitution floor,
// correction trigger, warm-start skip) are replayed from the forward's own
// values.  Friction enters through the per-contact coefficients mu1, mu2
// (mu_pair times a frozen direction scale), mass through the solver inverse
// mass and inverse inertia.
//
// Templated on the scalar so the same routines run on the device (float)
// and, for their checks, on the host (double).
#pragma once
#include <cmath>

#if defined(__HIPCC__)
#define VXA __host__ __device__ inline
#else
#define VXA inline
#endif

namespace vxadj {

template <class T>
struct V {
    T x, y, z;
};
template <class T> VXA V<T> vec(T x, T y, T z) { return V<T>{x, y, z}; }
template <class T> VXA V<T> operator+(V<T> a, V<T> b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
template <class T> VXA V<T> operator-(V<T> a, V<T> b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
template <class T> VXA V<T> operator*(V<T> a, T s) { return {a.x * s, a.y * s, a.z * s}; }
template <class T> VXA V<T>& operator+=(V<T>& a, V<T> b) { a.x += b.x; a.y += b.y; a.z += b.z; return a; }
template <class T> VXA V<T>& operator-=(V<T>& a, V<T> b) { a.x -= b.x; a.y -= b.y; a.z -= b.z; return a; }
template <class T> VXA T dot(V<T> a, V<T> b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
template <class T> VXA V<T> cross(V<T> a, V<T> b) {
    return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
}
template <class T> struct Q { T x, y, z, w; };

// Adjoint of c = a x b: a_bar += b x c_bar, b_bar += c_bar x a.
template <class T> VXA void crossAdj(V<T> a, V<T> b, V<T> cBar, V<T>& aBar, V<T>& bBar) {
    aBar += cross(b, cBar);
    bBar += cross(cBar, a);
}

// rotate(q, v) = v + 2 qv x (qv x v + w v) (math.h:113-118); q_bar from out_bar.
template <class T> VXA V<T> rotate(Q<T> q, V<T> v) {
    const V<T> qv{q.x, q.y, q.z};
    const V<T> t = cross(qv, cross(qv, v) + v * q.w);
    return v + t * T(2);
}
template <class T> VXA void rotateAdjQ(Q<T> q, V<T> v, V<T> outBar, Q<T>& qBar) {
    const V<T> qv{q.x, q.y, q.z};
    const V<T> u = cross(qv, v) + v * q.w;
    const V<T> tBar = outBar * T(2);
    V<T> qvBar{T(0), T(0), T(0)}, uBar{T(0), T(0), T(0)}, vBar{T(0), T(0), T(0)};
    crossAdj(qv, u, tBar, qvBar, uBar);
    crossAdj(qv, v, uBar, qvBar, vBar);
    qBar.x += qvBar.x;
    qBar.y += qvBar.y;
    qBar.z += qvBar.z;
    qBar.w += dot(v, uBar);
}

// One body side of a contact row: its solver inverse mass, isotropic inverse
// inertia, whether it is updated (dynamic), and its arm.
template <class T>
struct Side {
    T im, ii;
    bool dyn;
};

// Frozen and pose-derived data of one contact row for one substep.
template <class T>
struct Row {
    V<T> ra, rb, n, t1, t2;
    T kN, k1, k2, lg, h;  // effective masses, live gap, substep length
    T vn0, rest, mu1, mu2;
};

// Adjoints a row visit accumulates into its contact (summed over the visits
// of a substep, then turned into pose and parameter adjoints by rowPrefixAdj).
template <class T>
struct RowBar {
    V<T> ra, rb;
    T lg, kN, k1, k2, mu1, mu2, imA, imB, iiA, iiB;
};

template <class T> VXA void clear(RowBar<T>& b) {
    b.ra = b.rb = V<T>{T(0), T(0), T(0)};
    b.lg = b.kN = b.k1 = b.k2 = b.mu1 = b.mu2 = b.imA = b.imB = b.iiA = b.iiB = T(0);
}

// Applies impulse P to both sides: v_a += P m_a, w_a += I_a (r_a x P),
// v_b -= P m_b, w_b -= I_b (r_b x P).
template <class T>
VXA void applyImpulse(const Side<T>& A, const Side<T>& B, V<T> ra, V<T> rb, V<T> P, V<T>& vA, V<T>& wA, V<T>& vB,
                      V<T>& wB) {
    if (A.dyn) { vA += P * A.im; wA += cross(ra, P) * A.ii; }
    if (B.dyn) { vB -= P * B.im; wB -= cross(rb, P) * B.ii; }
}
// Its adjoint: P_bar and the side parameters' and arms' adjoints, given the
// velocity adjoints after the update (unchanged by it: the map is v + f(P)).
template <class T>
VXA V<T> applyImpulseAdj(const Side<T>& A, const Side<T>& B, V<T> ra, V<T> rb, V<T> P, V<T> vAb, V<T> wAb, V<T> vBb,
                         V<T> wBb, RowBar<T>& bar) {
    V<T> Pb{T(0), T(0), T(0)};
    if (A.dyn) {
        Pb += vAb * A.im;
        bar.imA += dot(P, vAb);
        const V<T> c = cross(ra, P);
        bar.iiA += dot(c, wAb);
        crossAdj(ra, P, wAb * A.ii, bar.ra, Pb);
    }
    if (B.dyn) {
        Pb -= vBb * B.im;
        bar.imB -= dot(P, vBb);
        const V<T> c = cross(rb, P);
        bar.iiB -= dot(c, wBb);
        crossAdj(rb, P, wBb * (-B.ii), bar.rb, Pb);
    }
    return Pb;
}

// Relative velocity of the contact points: (v_a + w_a x r_a) - (v_b + w_b x r_b).
template <class T> VXA V<T> relVel(V<T> vA, V<T> wA, V<T> vB, V<T> wB, V<T> ra, V<T> rb) {
    return (vA + cross(wA, ra)) - (vB + cross(wB, rb));
}
template <class T>
VXA void relVelAdj(V<T> wA, V<T> wB, V<T> ra, V<T> rb, V<T> rvBar, V<T>& vAb, V<T>& wAb, V<T>& vBb, V<T>& wBb,
                   RowBar<T>& bar) {
    vAb += rvBar;
    crossAdj(wA, ra, rvBar, wAb, bar.ra);
    const V<T> neg = rvBar * T(-1);
    vBb += neg;
    crossAdj(wB, rb, neg, wBb, bar.rb);
}

// State of one row visit: the two bodies' velocities and the accumulated
// normal and tangent impulses.
template <class T>
struct Visit {
    V<T> vA, wA, vB, wB;
    T cN, c1, c2;
};

// The row's forward (solveContact, two-axis path, no rolling or spinning
// resistance), in place.
template <class T>
VXA void rowForward(const Row<T>& r, const Side<T>& A, const Side<T>& B, Visit<T>& s) {
    const T vn = dot(relVel(s.vA, s.wA, s.vB, s.wB, r.ra, r.rb), r.n);
    T target = r.lg > T(0) ? -r.lg / r.h : T(0);
    if (r.vn0 < T(-1) && (r.lg <= T(0) || vn > T(0))) target = target > -r.rest * r.vn0 ? target : -r.rest * r.vn0;
    T jn = (target - vn) / r.kN;
    const T nextN = s.cN + jn > T(0) ? s.cN + jn : T(0);
    jn = nextN - s.cN;
    s.cN = nextN;
    applyImpulse(A, B, r.ra, r.rb, r.n * jn, s.vA, s.wA, s.vB, s.wB);
    if (r.k1 > T(0) && r.k2 > T(0)) {
        const V<T> rv = relVel(s.vA, s.wA, s.vB, s.wB, r.ra, r.rb);
        const T o1 = s.c1, o2 = s.c2;
        T n1 = o1 - dot(rv, r.t1) / r.k1;
        T n2 = o2 - dot(rv, r.t2) / r.k2;
        const T m1 = r.mu1 * s.cN, m2 = r.mu2 * s.cN;
        if (m1 <= T(0)) n1 = T(0);
        if (m2 <= T(0)) n2 = T(0);
        const T el = (m1 > T(0) ? n1 * n1 / (m1 * m1) : T(0)) + (m2 > T(0) ? n2 * n2 / (m2 * m2) : T(0));
        if (el > T(1)) {
            const T sc = T(1) / sqrt(el);
            n1 *= sc;
            n2 *= sc;
        }
        s.c1 = n1;
        s.c2 = n2;
        applyImpulse(A, B, r.ra, r.rb, r.t1 * (n1 - o1) + r.t2 * (n2 - o2), s.vA, s.wA, s.vB, s.wB);
    }
}

// The row's adjoint.  `pre` is the visit state before the row (the forward is
// recomputed from it); `bar` holds the adjoints of the state after the row on
// entry and of the state before it on return; the row's own data adjoints are
// added to `rb`.
template <class T>
VXA void rowAdjoint(const Row<T>& r, const Side<T>& A, const Side<T>& B, const Visit<T>& pre, Visit<T>& bar,
                    RowBar<T>& rbar) {
    // Forward recomputation, keeping what the reverse needs.
    const V<T> rv0 = relVel(pre.vA, pre.wA, pre.vB, pre.wB, r.ra, r.rb);
    const T vn = dot(rv0, r.n);
    const T base = r.lg > T(0) ? -r.lg / r.h : T(0);
    bool floorPicked = false;
    T target = base;
    if (r.vn0 < T(-1) && (r.lg <= T(0) || vn > T(0)) && !(base > -r.rest * r.vn0)) {
        target = -r.rest * r.vn0;
        floorPicked = true;
    }
    const T jn0 = (target - vn) / r.kN;
    const T sN = pre.cN + jn0;
    const bool clampN = !(sN > T(0));
    const T nN = clampN ? T(0) : sN;
    const T jn = nN - pre.cN;
    const V<T> P = r.n * jn;
    Visit<T> mid = pre;
    applyImpulse(A, B, r.ra, r.rb, P, mid.vA, mid.wA, mid.vB, mid.wB);
    mid.cN = nN;
    const bool friction = r.k1 > T(0) && r.k2 > T(0);
    // Reverse of the friction rows.
    T nNb = bar.cN;
    if (friction) {
        const V<T> rv = relVel(mid.vA, mid.wA, mid.vB, mid.wB, r.ra, r.rb);
        const T o1 = pre.c1, o2 = pre.c2;
        const T d1 = dot(rv, r.t1), d2 = dot(rv, r.t2);
        T n1 = o1 - d1 / r.k1, n2 = o2 - d2 / r.k2;
        const T m1 = r.mu1 * nN, m2 = r.mu2 * nN;
        const bool z1 = !(m1 > T(0)), z2 = !(m2 > T(0));
        if (z1) n1 = T(0);
        if (z2) n2 = T(0);
        const T el = (!z1 ? n1 * n1 / (m1 * m1) : T(0)) + (!z2 ? n2 * n2 / (m2 * m2) : T(0));
        const bool scaled = el > T(1);
        const T sc = scaled ? T(1) / sqrt(el) : T(1);
        const T f1 = n1 * sc, f2 = n2 * sc;
        const V<T> F = r.t1 * (f1 - o1) + r.t2 * (f2 - o2);
        const V<T> Fb = applyImpulseAdj(A, B, r.ra, r.rb, F, bar.vA, bar.wA, bar.vB, bar.wB, rbar);
        T f1b = bar.c1 + dot(r.t1, Fb), f2b = bar.c2 + dot(r.t2, Fb);
        T o1b = -dot(r.t1, Fb), o2b = -dot(r.t2, Fb);
        T n1b = f1b * sc, n2b = f2b * sc, elb = T(0);
        if (scaled) elb = T(-0.5) * sc * sc * sc * (n1 * f1b + n2 * f2b);
        T m1b = T(0), m2b = T(0);
        if (!z1) {
            n1b += elb * T(2) * n1 / (m1 * m1);
            m1b = -elb * T(2) * n1 * n1 / (m1 * m1 * m1);
        } else {
            n1b = T(0);
        }
        if (!z2) {
            n2b += elb * T(2) * n2 / (m2 * m2);
            m2b = -elb * T(2) * n2 * n2 / (m2 * m2 * m2);
        } else {
            n2b = T(0);
        }
        rbar.mu1 += m1b * nN;
        rbar.mu2 += m2b * nN;
        nNb += m1b * r.mu1 + m2b * r.mu2;
        o1b += n1b;
        o2b += n2b;
        const V<T> rvb = r.t1 * (-n1b / r.k1) + r.t2 * (-n2b / r.k2);
        rbar.k1 += n1b * d1 / (r.k1 * r.k1);
        rbar.k2 += n2b * d2 / (r.k2 * r.k2);
        relVelAdj(mid.wA, mid.wB, r.ra, r.rb, rvb, bar.vA, bar.wA, bar.vB, bar.wB, rbar);
        bar.c1 = o1b;
        bar.c2 = o2b;
    }
    // Reverse of the normal row.
    const V<T> Pb = applyImpulseAdj(A, B, r.ra, r.rb, P, bar.vA, bar.wA, bar.vB, bar.wB, rbar);
    const T jnb = dot(r.n, Pb);
    nNb += jnb;
    T cNb = -jnb;
    const T sNb = clampN ? T(0) : nNb;
    cNb += sNb;
    const T jn0b = sNb;
    const T targetb = jn0b / r.kN;
    const T vnb = -jn0b / r.kN;
    rbar.kN += -jn0b * jn0 / r.kN;
    if (!floorPicked && r.lg > T(0)) rbar.lg += -targetb / r.h;
    relVelAdj(pre.wA, pre.wB, r.ra, r.rb, r.n * vnb, bar.vA, bar.wA, bar.vB, bar.wB, rbar);
    bar.cN = cNb;
}

// warmStartContact: the stored impulses applied once (skipped when all are
// zero and the normal one is not positive).
template <class T>
VXA bool warmSkips(T cN, T c1, T c2) { return cN <= T(0) && c1 == T(0) && c2 == T(0); }
template <class T>
VXA void warmForward(const Row<T>& r, const Side<T>& A, const Side<T>& B, Visit<T>& s) {
    if (warmSkips(s.cN, s.c1, s.c2)) return;
    applyImpulse(A, B, r.ra, r.rb, r.n * s.cN + r.t1 * s.c1 + r.t2 * s.c2, s.vA, s.wA, s.vB, s.wB);
}
// Adds the impulses' adjoints to bar.cN/c1/c2 (velocity adjoints pass through).
template <class T>
VXA void warmAdjoint(const Row<T>& r, const Side<T>& A, const Side<T>& B, const Visit<T>& pre, Visit<T>& bar,
                     RowBar<T>& rbar) {
    if (warmSkips(pre.cN, pre.c1, pre.c2)) return;
    const V<T> P = r.n * pre.cN + r.t1 * pre.c1 + r.t2 * pre.c2;
    const V<T> Pb = applyImpulseAdj(A, B, r.ra, r.rb, P, bar.vA, bar.wA, bar.vB, bar.wB, rbar);
    bar.cN += dot(r.n, Pb);
    bar.c1 += dot(r.t1, Pb);
    bar.c2 += dot(r.t2, Pb);
}

// Row prefix: k = m_a + m_b + I_a |r_a x d|^2 + I_b |r_b x d|^2 for d = n, t1,
// t2, and the live gap from the arms.  Turns a contact's accumulated RowBar
// into arm and side-parameter adjoints and adds the live gap's share to the
// positions' adjoints (the arms then go to the orientations, rotateAdjQ).
template <class T> VXA V<T> armSqGrad(V<T> a, V<T> d) {  // d/da |a x d|^2
    return (a * dot(d, d) - d * dot(a, d)) * T(2);
}
template <class T>
VXA void rowPrefixAdj(const Row<T>& r, const Side<T>& A, const Side<T>& B, RowBar<T>& b, V<T>& xAb, V<T>& xBb) {
    const T ks[3] = {b.kN, b.k1, b.k2};
    const V<T> ds[3] = {r.n, r.t1, r.t2};
    for (int q = 0; q < 3; ++q) {
        b.imA += ks[q];
        b.imB += ks[q];
        if (A.dyn) {
            const V<T> c = cross(r.ra, ds[q]);
            b.iiA += ks[q] * dot(c, c);
            b.ra += armSqGrad(r.ra, ds[q]) * (ks[q] * A.ii);
        }
        if (B.dyn) {
            const V<T> c = cross(r.rb, ds[q]);
            b.iiB += ks[q] * dot(c, c);
            b.rb += armSqGrad(r.rb, ds[q]) * (ks[q] * B.ii);
        }
    }
    // lg = bias0 + n.((x_a + r_a) - (x_b + r_b)).
    b.ra += r.n * b.lg;
    b.rb -= r.n * b.lg;
    xAb += r.n * b.lg;
    xBb -= r.n * b.lg;
}

// Transform advance of a dynamic body with isotropic inertia over h:
// x' = x + v h, q' = normalize(q + (w h / 2, 0) q).  Given x'_bar, q'_bar,
// adds to x_bar, q_bar, v_bar, w_bar.
template <class T>
VXA Q<T> quatStep(Q<T> q, V<T> w, T h) {
    const T ax = w.x * T(0.5) * h, ay = w.y * T(0.5) * h, az = w.z * T(0.5) * h;
    Q<T> u{q.x + (ax * q.w + ay * q.z - az * q.y), q.y + (-ax * q.z + ay * q.w + az * q.x),
           q.z + (ax * q.y - ay * q.x + az * q.w), q.w + (-ax * q.x - ay * q.y - az * q.z)};
    const T inv = T(1) / sqrt(u.x * u.x + u.y * u.y + u.z * u.z + u.w * u.w);
    return {u.x * inv, u.y * inv, u.z * inv, u.w * inv};
}
template <class T>
VXA void transformAdj(Q<T> q, V<T> w, T h, V<T> xb1, Q<T> qb1, V<T>& xb, Q<T>& qb, V<T>& vb, V<T>& wb) {
    xb += xb1;
    vb += xb1 * h;
    const T ax = w.x * T(0.5) * h, ay = w.y * T(0.5) * h, az = w.z * T(0.5) * h;
    const Q<T> u{q.x + (ax * q.w + ay * q.z - az * q.y), q.y + (-ax * q.z + ay * q.w + az * q.x),
                 q.z + (ax * q.y - ay * q.x + az * q.w), q.w + (-ax * q.x - ay * q.y - az * q.z)};
    const T len = sqrt(u.x * u.x + u.y * u.y + u.z * u.z + u.w * u.w);
    const Q<T> o{u.x / len, u.y / len, u.z / len, u.w / len};
    const T proj = o.x * qb1.x + o.y * qb1.y + o.z * qb1.z + o.w * qb1.w;
    const Q<T> ub{(qb1.x - o.x * proj) / len, (qb1.y - o.y * proj) / len, (qb1.z - o.z * proj) / len,
                  (qb1.w - o.w * proj) / len};
    qb.x += ub.x + (az * ub.y - ay * ub.z - ax * ub.w);
    qb.y += ub.y + (-az * ub.x + ax * ub.z - ay * ub.w);
    qb.z += ub.z + (ay * ub.x - ax * ub.y - az * ub.w);
    qb.w += ub.w + (ax * ub.x + ay * ub.y + az * ub.z);
    const T abx = q.w * ub.x - q.z * ub.y + q.y * ub.z - q.x * ub.w;
    const T aby = q.z * ub.x + q.w * ub.y - q.x * ub.z - q.y * ub.w;
    const T abz = -q.y * ub.x + q.x * ub.y + q.w * ub.z - q.z * ub.w;
    wb += V<T>{abx, aby, abz} * (T(0.5) * h);
}

// q / |q| at the start of a step (the orientation renormalisation); adjoint.
template <class T>
VXA void normalizeAdj(Q<T> q, Q<T>& qb) {
    const T len = sqrt(q.x * q.x + q.y * q.y + q.z * q.z + q.w * q.w);
    const Q<T> o{q.x / len, q.y / len, q.z / len, q.w / len};
    const T proj = o.x * qb.x + o.y * qb.y + o.z * qb.z + o.w * qb.w;
    qb = Q<T>{(qb.x - o.x * proj) / len, (qb.y - o.y * proj) / len, (qb.z - o.z * proj) / len,
              (qb.w - o.w * proj) / len};
}

// Positional correction of one contact (world.cpp:4371-4394, no stack
// stabilisation): g = bias0 + n.(pa - pb); when g < -1e-3 the bodies move
// apart along n by -0.6 (g + 1e-3) / (m_a + m_b) times their inverse mass.
// Returns whether it fired; `push` receives the push.
template <class T>
VXA bool correctionFires(T g, T imA, T imB) { return imA + imB > T(0) && g < T(-1e-3); }
template <class T>
VXA T correctionPush(T g, T imA, T imB) { return T(-0.6) * (g + T(1e-3)) / (imA + imB); }
// Adjoint of a fired correction, given the positions' adjoints after it (the
// map is x + f): returns g_bar; adds the inverse masses' adjoints.
template <class T>
VXA T correctionAdj(T g, T imA, T imB, bool dynA, bool dynB, V<T> n, V<T> xAb, V<T> xBb, T& imAb, T& imBb) {
    const T push = correctionPush(g, imA, imB);
    T pushb = T(0);
    if (dynA) { pushb += imA * dot(n, xAb); imAb += push * dot(n, xAb); }
    if (dynB) { pushb -= imB * dot(n, xBb); imBb -= push * dot(n, xBb); }
    const T S = imA + imB;
    const T Sb = T(0.6) * (g + T(1e-3)) / (S * S) * pushb;
    imAb += Sb;
    imBb += Sb;
    return T(-0.6) / S * pushb;
}

}  // namespace vxadj
