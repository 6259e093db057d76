// Checks the hand-written adjoints of velox_amd/csrc/vxd_adjoint.hpp against
// central finite differences of their own forward maps, in double precision,
// on seeded random cases: a contact row (arms and live gap from the poses,
// effective masses, the normal and both friction rows), the warm start, the
// transform advance, the start-of-step renormalisation and the positional
// correction.  Every input of each map is perturbed; a case whose branch
// pattern changes within the step is skipped.  Prints the worst relative
// error per map and exits non-zero above the bound.
#include "../velox_amd/csrc/vxd_adjoint.hpp"
#include <cmath>
#include <cstdio>
#include <random>
#include <vector>

using namespace vxadj;
using D = double;

namespace {
std::mt19937_64 rng(2024);
D uni(D lo, D hi) { return std::uniform_real_distribution<D>(lo, hi)(rng); }
V<D> rv3(D s) { return {uni(-s, s), uni(-s, s), uni(-s, s)}; }
Q<D> rq() {
    Q<D> q{uni(-1, 1), uni(-1, 1), uni(-1, 1), uni(-1, 1)};
    const D l = std::sqrt(q.x * q.x + q.y * q.y + q.z * q.z + q.w * q.w);
    return {q.x / l, q.y / l, q.z / l, q.w / l};
}
V<D> unit(V<D> v) { const D l = std::sqrt(dot(v, v)); return v * (1.0 / l); }
void tangents(V<D> n, V<D>& t1, V<D>& t2) {  // narrowphase.h:38-50 (frozen with the normal)
    const D ax = std::fabs(n.x), ay = std::fabs(n.y), az = std::fabs(n.z);
    const V<D> ref = ax < ay ? (ax < az ? V<D>{1, 0, 0} : V<D>{0, 0, 1}) : (ay < az ? V<D>{0, 1, 0} : V<D>{0, 0, 1});
    t1 = unit(cross(n, ref));
    t2 = cross(n, t1);
}

// Row inputs as a flat vector: xA(3) qA(4) xB(3) qB(4) vA wA vB wB (12) cN c1 c2
// imA iiA imB iiB mu1 mu2.
constexpr int kIn = 3 + 4 + 3 + 4 + 12 + 3 + 4 + 2;
struct RowCase {
    V<D> lA, lB, n, t1, t2;
    D bias0, h, vn0, rest;
    bool dynB;
};

struct Unpacked {
    V<D> xA, xB, vA, wA, vB, wB;
    Q<D> qA, qB;
    D cN, c1, c2, imA, iiA, imB, iiB, mu1, mu2;
};
Unpacked unpack(const std::vector<D>& p) {
    Unpacked u;
    int k = 0;
    auto v = [&]() { V<D> r{p[k], p[k + 1], p[k + 2]}; k += 3; return r; };
    auto q = [&]() { Q<D> r{p[k], p[k + 1], p[k + 2], p[k + 3]}; k += 4; return r; };
    u.xA = v(); u.qA = q(); u.xB = v(); u.qB = q();
    u.vA = v(); u.wA = v(); u.vB = v(); u.wB = v();
    u.cN = p[k++]; u.c1 = p[k++]; u.c2 = p[k++];
    u.imA = p[k++]; u.iiA = p[k++]; u.imB = p[k++]; u.iiB = p[k++]; u.mu1 = p[k++]; u.mu2 = p[k++];
    return u;
}

void makeRow(const RowCase& c, const Unpacked& u, Row<D>& r, Side<D>& A, Side<D>& B) {
    A = {u.imA, u.iiA, true};
    B = {c.dynB ? u.imB : 0.0, c.dynB ? u.iiB : 0.0, c.dynB};
    r.ra = rotate(u.qA, c.lA);
    r.rb = rotate(u.qB, c.lB);
    r.n = c.n; r.t1 = c.t1; r.t2 = c.t2;
    r.lg = c.bias0 + dot(c.n, (u.xA + r.ra) - (u.xB + r.rb));
    auto k = [&](V<D> d) {
        const V<D> ca = cross(r.ra, d), cb = cross(r.rb, d);
        return A.im + B.im + A.ii * dot(ca, ca) + (B.dyn ? B.ii * dot(cb, cb) : 0.0);
    };
    r.kN = k(c.n); r.k1 = k(c.t1); r.k2 = k(c.t2);
    r.h = c.h; r.vn0 = c.vn0; r.rest = c.rest; r.mu1 = u.mu1; r.mu2 = u.mu2;
}

// Branch pattern of the row at these inputs (to skip cases that straddle one).
int rowBranches(const Row<D>& r, const Visit<D>& s0) {
    Visit<D> s = s0;
    const D vn = dot(relVel(s.vA, s.wA, s.vB, s.wB, r.ra, r.rb), r.n);
    int b = (r.lg > 0) | ((r.vn0 < -1 && (r.lg <= 0 || vn > 0)) << 1);
    const D base = r.lg > 0 ? -r.lg / r.h : 0.0;
    b |= (base > -r.rest * r.vn0) << 2;
    const D target = (b & 2) && !(b & 4) ? -r.rest * r.vn0 : base;
    const D sN = s.cN + (target - vn) / r.kN;
    b |= (sN > 0) << 3;
    const D nN = sN > 0 ? sN : 0.0;
    b |= (r.mu1 * nN > 0) << 4;
    b |= (r.mu2 * nN > 0) << 5;
    return b;
}

// The ellipse scaling's branch, from a full forward pass.
bool rowScales(const Row<D>& r, const Side<D>& A, const Side<D>& B, Visit<D> s) {
    Visit<D> t = s;
    rowForward(r, A, B, t);
    const D m1 = r.mu1 * t.cN, m2 = r.mu2 * t.cN;
    // scaled exactly when the kept tangent impulses sit on the ellipse
    const D el = (m1 > 0 ? t.c1 * t.c1 / (m1 * m1) : 0.0) + (m2 > 0 ? t.c2 * t.c2 / (m2 * m2) : 0.0);
    return std::fabs(el - 1.0) < 1e-9;
}

D rowLoss(const RowCase& c, const std::vector<D>& p, const std::vector<D>& w, int* branches) {
    const Unpacked u = unpack(p);
    Row<D> r; Side<D> A, B;
    makeRow(c, u, r, A, B);
    Visit<D> s{u.vA, u.wA, u.vB, u.wB, u.cN, u.c1, u.c2};
    if (branches) *branches = rowBranches(r, s) | (rowScales(r, A, B, s) << 6);
    rowForward(r, A, B, s);
    const D out[15] = {s.vA.x, s.vA.y, s.vA.z, s.wA.x, s.wA.y, s.wA.z, s.vB.x, s.vB.y, s.vB.z,
                       s.wB.x, s.wB.y, s.wB.z, s.cN, s.c1, s.c2};
    D L = 0;
    for (int i = 0; i < 15; ++i) L += w[i] * out[i];
    return L;
}

std::vector<D> rowGrad(const RowCase& c, const std::vector<D>& p, const std::vector<D>& w) {
    const Unpacked u = unpack(p);
    Row<D> r; Side<D> A, B;
    makeRow(c, u, r, A, B);
    const Visit<D> pre{u.vA, u.wA, u.vB, u.wB, u.cN, u.c1, u.c2};
    Visit<D> bar{{w[0], w[1], w[2]}, {w[3], w[4], w[5]}, {w[6], w[7], w[8]}, {w[9], w[10], w[11]}, w[12], w[13], w[14]};
    RowBar<D> rb;
    clear(rb);
    rowAdjoint(r, A, B, pre, bar, rb);
    V<D> xAb{0, 0, 0}, xBb{0, 0, 0};
    rowPrefixAdj(r, A, B, rb, xAb, xBb);
    Q<D> qAb{0, 0, 0, 0}, qBb{0, 0, 0, 0};
    rotateAdjQ(u.qA, c.lA, rb.ra, qAb);
    rotateAdjQ(u.qB, c.lB, rb.rb, qBb);
    std::vector<D> g = {xAb.x, xAb.y, xAb.z, qAb.x, qAb.y, qAb.z, qAb.w, xBb.x, xBb.y, xBb.z, qBb.x, qBb.y, qBb.z,
                        qBb.w, bar.vA.x, bar.vA.y, bar.vA.z, bar.wA.x, bar.wA.y, bar.wA.z, bar.vB.x, bar.vB.y,
                        bar.vB.z, bar.wB.x, bar.wB.y, bar.wB.z, bar.cN, bar.c1, bar.c2, rb.imA, rb.iiA,
                        c.dynB ? rb.imB : 0.0, c.dynB ? rb.iiB : 0.0, rb.mu1, rb.mu2};
    return g;
}

D worstRel(const std::vector<D>& a, const std::vector<D>& b) {
    D num = 0, den = 0;
    for (size_t i = 0; i < a.size(); ++i) { num += (a[i] - b[i]) * (a[i] - b[i]); den += b[i] * b[i]; }
    return std::sqrt(num) / (std::sqrt(den) + 1e-12);
}

D checkRows(int cases, int& used) {
    D worst = 0;
    used = 0;
    for (int it = 0; it < cases; ++it) {
        RowCase c;
        c.lA = rv3(0.5); c.lB = rv3(0.5);
        c.n = unit(rv3(1.0));
        tangents(c.n, c.t1, c.t2);
        c.bias0 = uni(-0.02, 0.02);
        c.h = 1.0 / 240.0;
        c.vn0 = it % 3 == 0 ? uni(-4, -1.5) : uni(-0.5, 0.5);
        c.rest = it % 2 ? 0.0 : uni(0, 0.5);
        c.dynB = it % 4 != 0;
        std::vector<D> p(kIn);
        int k = 0;
        auto put3 = [&](V<D> v) { p[k++] = v.x; p[k++] = v.y; p[k++] = v.z; };
        auto put4 = [&](Q<D> q) { p[k++] = q.x; p[k++] = q.y; p[k++] = q.z; p[k++] = q.w; };
        put3(rv3(0.3)); put4(rq()); put3(rv3(0.3)); put4(rq());
        for (int j = 0; j < 4; ++j) put3(rv3(it % 5 == 0 ? 3.0 : 0.6));
        p[k++] = it % 7 == 0 ? 0.0 : uni(0, 3); p[k++] = uni(-1, 1); p[k++] = uni(-1, 1);
        p[k++] = uni(0.2, 2.0); p[k++] = uni(0.5, 20.0); p[k++] = uni(0.2, 2.0); p[k++] = uni(0.5, 20.0);
        p[k++] = it % 11 == 0 ? 0.0 : uni(0.1, 1.0); p[k++] = uni(0.1, 1.0);
        std::vector<D> w(15);
        for (D& x : w) x = uni(-1, 1);
        const std::vector<D> g = rowGrad(c, p, w);
        std::vector<D> fd(kIn);
        const D h = 1e-6;
        bool straddles = false;
        int b0 = 0;
        rowLoss(c, p, w, &b0);
        for (int i = 0; i < kIn; ++i) {
            std::vector<D> pp = p, pm = p;
            pp[i] += h; pm[i] -= h;
            int bp = 0, bm = 0;
            fd[i] = (rowLoss(c, pp, w, &bp) - rowLoss(c, pm, w, &bm)) / (2 * h);
            straddles |= bp != b0 || bm != b0;
        }
        if (straddles) continue;
        ++used;
        const D e = worstRel(g, fd);
        if (e > worst) worst = e;
    }
    return worst;
}

D checkWarm(int cases) {
    D worst = 0;
    for (int it = 0; it < cases; ++it) {
        const Q<D> qA = rq(), qB = rq();
        const V<D> lA = rv3(0.5), lB = rv3(0.5), n = unit(rv3(1));
        V<D> t1, t2;
        tangents(n, t1, t2);
        const bool dynB = it % 3 != 0;
        auto loss = [&](const std::vector<D>& p, const std::vector<D>& w) {
            Row<D> r{};
            r.ra = rotate(Q<D>{p[0], p[1], p[2], p[3]}, lA);
            r.rb = rotate(Q<D>{p[4], p[5], p[6], p[7]}, lB);
            r.n = n; r.t1 = t1; r.t2 = t2;
            Side<D> A{p[8], p[9], true}, B{dynB ? p[10] : 0.0, dynB ? p[11] : 0.0, dynB};
            Visit<D> s{{p[12], p[13], p[14]}, {p[15], p[16], p[17]}, {p[18], p[19], p[20]}, {p[21], p[22], p[23]},
                       p[24], p[25], p[26]};
            warmForward(r, A, B, s);
            const D out[12] = {s.vA.x, s.vA.y, s.vA.z, s.wA.x, s.wA.y, s.wA.z,
                               s.vB.x, s.vB.y, s.vB.z, s.wB.x, s.wB.y, s.wB.z};
            D L = 0;
            for (int i = 0; i < 12; ++i) L += w[i] * out[i];
            return L;
        };
        std::vector<D> p = {qA.x, qA.y, qA.z, qA.w, qB.x, qB.y, qB.z, qB.w, uni(0.2, 2), uni(1, 10), uni(0.2, 2),
                            uni(1, 10)};
        for (int j = 0; j < 12; ++j) p.push_back(uni(-1, 1));
        p.push_back(uni(0.1, 2)); p.push_back(uni(-1, 1)); p.push_back(uni(-1, 1));
        std::vector<D> w(12);
        for (D& x : w) x = uni(-1, 1);
        Row<D> r{};
        r.ra = rotate(qA, lA); r.rb = rotate(qB, lB); r.n = n; r.t1 = t1; r.t2 = t2;
        Side<D> A{p[8], p[9], true}, B{dynB ? p[10] : 0.0, dynB ? p[11] : 0.0, dynB};
        const Visit<D> pre{{p[12], p[13], p[14]}, {p[15], p[16], p[17]}, {p[18], p[19], p[20]}, {p[21], p[22], p[23]},
                           p[24], p[25], p[26]};
        Visit<D> bar{{w[0], w[1], w[2]}, {w[3], w[4], w[5]}, {w[6], w[7], w[8]}, {w[9], w[10], w[11]}, 0, 0, 0};
        RowBar<D> rb;
        clear(rb);
        warmAdjoint(r, A, B, pre, bar, rb);
        Q<D> qAb{0, 0, 0, 0}, qBb{0, 0, 0, 0};
        rotateAdjQ(qA, lA, rb.ra, qAb);
        rotateAdjQ(qB, lB, rb.rb, qBb);
        std::vector<D> g = {qAb.x, qAb.y, qAb.z, qAb.w, qBb.x, qBb.y, qBb.z, qBb.w, rb.imA, rb.iiA,
                            dynB ? rb.imB : 0.0, dynB ? rb.iiB : 0.0,
                            bar.vA.x, bar.vA.y, bar.vA.z, bar.wA.x, bar.wA.y, bar.wA.z, bar.vB.x, bar.vB.y, bar.vB.z,
                            bar.wB.x, bar.wB.y, bar.wB.z, bar.cN, bar.c1, bar.c2};
        std::vector<D> fd(p.size());
        for (size_t i = 0; i < p.size(); ++i) {
            std::vector<D> pp = p, pm = p;
            pp[i] += 1e-6; pm[i] -= 1e-6;
            fd[i] = (loss(pp, w) - loss(pm, w)) / 2e-6;
        }
        const D e = worstRel(g, fd);
        if (e > worst) worst = e;
    }
    return worst;
}

D checkTransform(int cases) {
    D worst = 0;
    for (int it = 0; it < cases; ++it) {
        const D h = 1.0 / 240.0;
        std::vector<D> w(7);
        for (D& x : w) x = uni(-1, 1);
        const Q<D> q0 = rq();
        std::vector<D> p = {uni(-1, 1), uni(-1, 1), uni(-1, 1), q0.x, q0.y, q0.z, q0.w,
                            uni(-3, 3), uni(-3, 3), uni(-3, 3), uni(-30, 30), uni(-30, 30), uni(-30, 30)};
        auto loss = [&](const std::vector<D>& v) {
            const V<D> x{v[0], v[1], v[2]}, vel{v[7], v[8], v[9]}, om{v[10], v[11], v[12]};
            const Q<D> q{v[3], v[4], v[5], v[6]};
            const V<D> x1 = x + vel * h;
            const Q<D> q1 = quatStep(q, om, h);
            return w[0] * x1.x + w[1] * x1.y + w[2] * x1.z + w[3] * q1.x + w[4] * q1.y + w[5] * q1.z + w[6] * q1.w;
        };
        V<D> xb{0, 0, 0}, vb{0, 0, 0}, wb{0, 0, 0};
        Q<D> qb{0, 0, 0, 0};
        transformAdj(Q<D>{p[3], p[4], p[5], p[6]}, V<D>{p[10], p[11], p[12]}, h, V<D>{w[0], w[1], w[2]},
                     Q<D>{w[3], w[4], w[5], w[6]}, xb, qb, vb, wb);
        const std::vector<D> g = {xb.x, xb.y, xb.z, qb.x, qb.y, qb.z, qb.w, vb.x, vb.y, vb.z, wb.x, wb.y, wb.z};
        std::vector<D> fd(p.size());
        for (size_t i = 0; i < p.size(); ++i) {
            std::vector<D> pp = p, pm = p;
            pp[i] += 1e-6; pm[i] -= 1e-6;
            fd[i] = (loss(pp) - loss(pm)) / 2e-6;
        }
        const D e = worstRel(g, fd);
        if (e > worst) worst = e;
        // renormalisation
        Q<D> q{uni(-1, 1), uni(-1, 1), uni(-1, 1), uni(-1, 1)};
        Q<D> qbn{w[3], w[4], w[5], w[6]};
        auto nloss = [&](Q<D> v) {
            const D l = std::sqrt(v.x * v.x + v.y * v.y + v.z * v.z + v.w * v.w);
            return (w[3] * v.x + w[4] * v.y + w[5] * v.z + w[6] * v.w) / l;
        };
        normalizeAdj(q, qbn);
        D qs[4] = {q.x, q.y, q.z, q.w}, gn[4] = {qbn.x, qbn.y, qbn.z, qbn.w};
        std::vector<D> a(4), b(4);
        for (int i = 0; i < 4; ++i) {
            D pp[4] = {qs[0], qs[1], qs[2], qs[3]}, pm[4] = {qs[0], qs[1], qs[2], qs[3]};
            pp[i] += 1e-6; pm[i] -= 1e-6;
            b[i] = (nloss({pp[0], pp[1], pp[2], pp[3]}) - nloss({pm[0], pm[1], pm[2], pm[3]})) / 2e-6;
            a[i] = gn[i];
        }
        const D en = worstRel(a, b);
        if (en > worst) worst = en;
    }
    return worst;
}

D checkCorrection(int cases, int& used) {
    D worst = 0;
    used = 0;
    for (int it = 0; it < cases; ++it) {
        const V<D> lA = rv3(0.5), lB = rv3(0.5), n = unit(rv3(1));
        const bool dynB = it % 3 != 0;
        const D bias0 = uni(-0.05, -0.002);
        std::vector<D> w(6);
        for (D& x : w) x = uni(-1, 1);
        const Q<D> qA = rq(), qB = rq();
        std::vector<D> p = {uni(-0.01, 0.01), uni(-0.01, 0.01), uni(-0.01, 0.01), qA.x, qA.y, qA.z, qA.w,
                            uni(-0.01, 0.01), uni(-0.01, 0.01), uni(-0.01, 0.01), qB.x, qB.y, qB.z, qB.w,
                            uni(0.2, 2), uni(0.2, 2)};
        auto run = [&](const std::vector<D>& v, bool* fired) {
            V<D> xA{v[0], v[1], v[2]}, xB{v[7], v[8], v[9]};
            const Q<D> qa{v[3], v[4], v[5], v[6]}, qb{v[10], v[11], v[12], v[13]};
            const D imA = v[14], imB = dynB ? v[15] : 0.0;
            const D g = bias0 + dot(n, (xA + rotate(qa, lA)) - (xB + rotate(qb, lB)));
            *fired = correctionFires(g, imA, imB);
            if (*fired) {
                const D push = correctionPush(g, imA, imB);
                xA += n * (push * imA);
                if (dynB) xB -= n * (push * imB);
            }
            return w[0] * xA.x + w[1] * xA.y + w[2] * xA.z + w[3] * xB.x + w[4] * xB.y + w[5] * xB.z;
        };
        bool f0 = false;
        run(p, &f0);
        if (!f0) continue;
        V<D> xAb{w[0], w[1], w[2]}, xBb{w[3], w[4], w[5]};
        D imAb = 0, imBb = 0;
        const V<D> xA{p[0], p[1], p[2]}, xB{p[7], p[8], p[9]};
        const Q<D> qa{p[3], p[4], p[5], p[6]}, qb{p[10], p[11], p[12], p[13]};
        const D g = bias0 + dot(n, (xA + rotate(qa, lA)) - (xB + rotate(qb, lB)));
        const D gb = correctionAdj(g, p[14], dynB ? p[15] : 0.0, true, dynB, n, xAb, xBb, imAb, imBb);
        Q<D> qAb{0, 0, 0, 0}, qBb{0, 0, 0, 0};
        V<D> gA = xAb + n * gb, gB = xBb - n * gb;
        rotateAdjQ(qa, lA, n * gb, qAb);
        rotateAdjQ(qb, lB, n * (-gb), qBb);
        const std::vector<D> grad = {gA.x, gA.y, gA.z, qAb.x, qAb.y, qAb.z, qAb.w, gB.x, gB.y, gB.z,
                                     qBb.x, qBb.y, qBb.z, qBb.w, imAb, dynB ? imBb : 0.0};
        std::vector<D> fd(p.size());
        bool straddles = false;
        for (size_t i = 0; i < p.size(); ++i) {
            std::vector<D> pp = p, pm = p;
            pp[i] += 1e-7; pm[i] -= 1e-7;
            bool a = false, b = false;
            fd[i] = (run(pp, &a) - run(pm, &b)) / 2e-7;
            straddles |= !a || !b;
        }
        if (straddles) continue;
        ++used;
        const D e = worstRel(grad, fd);
        if (e > worst) worst = e;
    }
    return worst;
}
}  // namespace

int main() {
    int rowsUsed = 0, corrUsed = 0;
    const D er = checkRows(4000, rowsUsed);
    const D ew = checkWarm(500);
    const D et = checkTransform(500);
    const D ec = checkCorrection(1500, corrUsed);
    std::printf("row %.3e (%d cases) warm %.3e transform %.3e correction %.3e (%d cases)\n", er, rowsUsed, ew, et, ec,
                corrUsed);
    const bool ok = er < 1e-6 && ew < 1e-6 && et < 1e-6 && ec < 1e-6 && rowsUsed > 2000 && corrUsed > 200;
    std::printf("%s\n", ok ? "adjoints ok" : "adjoint mismatch");
    return ok ? 0 : 1;
}
