// ORACLE -- test infrastructure only.  Only tests/, __graft_entry__.smoke() and
// bench.py's cpu_baseline leg may load liboracle.so, and only as the checker.
//
// Sequential CPU restatement of the device step that libvxhip runs on the GPU:
//   integrate -> exact-AABB broad phase -> narrow phase in canonical order ->
//   wake pass -> warm-start matching -> greedy coloring -> colored TGS solve
//   over substeps -> conservative-advancement net -> positional correction.
// Reference semantics followed (file:line in the Velox sources):
//   step order            src/world.cpp:3830-4397
//   broad-phase pair set  src/backend_cuda.cu:174-250 (exact AABB, filters, inert skip)
//   canonical order       src/backend_cuda.cu:587-595 (sort by (a<<32|b, point.xyz))
//   coloring              src/backend_cuda.cu:597-640
//   colored sweeps        src/backend_cuda.cu:652-689, 730-747 (2*I sweeps per substep)
#include "vxo_solver.hpp"
#include <algorithm>
#include <cstring>
#include <cstdio>
#include <cstdlib>
#include <vector>

using namespace vxo;

namespace {

Soup soupOf(const vx_soup* s) {
    Soup o;
    if (!s) return o;
    o.vertices = s->vertices; o.indices = s->indices; o.meshes = s->meshes; o.nodes = s->bvhNodes;
    o.triRefs = s->bvhTriRefs; o.hullPoints = s->hullPoints; o.hullFaceIndices = s->hullFaceIndices;
    o.children = s->compoundChildren;
    return o;
}

// Contacts of a plane/mesh pair name the convex body first
// (narrowphase.h:665-671), so a pair's contact key can be swapped.
inline uint64_t contactKeyOfPair(const vx_body* bodies, uint32_t i, uint32_t j) {
    const uint8_t si = bodies[i].shape;
    if (si == VX_SHAPE_PLANE || si == VX_SHAPE_MESH) return (uint64_t)j << 32 | i;
    return (uint64_t)i << 32 | j;
}

// backend_cuda.cu:587-595 comparator
inline bool canonicalLess(const vx_contact& x, const vx_contact& y) {
    uint64_t kx = (uint64_t)x.a << 32 | x.b, ky = (uint64_t)y.a << 32 | y.b;
    if (kx != ky) return kx < ky;
    if (x.point.x != y.point.x) return x.point.x < y.point.x;
    if (x.point.y != y.point.y) return x.point.y < y.point.y;
    return x.point.z < y.point.z;
}

void aabbOf(const vx_body& b, const Soup& soup, float dt, V3& lo, V3& hi) {
    if (b.shape == VX_SHAPE_PLANE) { lo = {-1e30f, -1e30f, -1e30f}; hi = {1e30f, 1e30f, 1e30f}; return; }
    if (b.shape == VX_SHAPE_MESH) {
        const vx_mesh& m = soup.meshes[b.meshIndex];
        lo = v3(m.aabbMin); hi = v3(m.aabbMax); return;
    }
    bodyAabb(b, dt, lo, hi);
}

// Candidate pairs {i<j : AABB overlap, filter passes, not both inert}, sorted by
// contact key.  O(N^2): this is the brute-force definition of the set that the
// device grid (backend_cuda.cu:174-250) must reproduce exactly.
std::vector<uint64_t> broadphase(const vx_body* bodies, uint32_t n, const Soup& soup, float dt) {
    std::vector<V3> lo(n), hi(n);
    for (uint32_t i = 0; i < n; ++i) aabbOf(bodies[i], soup, dt, lo[i], hi[i]);
    std::vector<uint64_t> pairs;
    for (uint32_t i = 0; i < n; ++i)
        for (uint32_t j = i + 1; j < n; ++j) {
            if (!aabbOverlap(lo[i], hi[i], lo[j], hi[j])) continue;
            if (inert(bodies[i]) && inert(bodies[j])) continue;
            if (!canCollide(bodies[i], bodies[j])) continue;
            pairs.push_back(contactKeyOfPair(bodies, i, j));
        }
    std::sort(pairs.begin(), pairs.end());
    return pairs;
}

void narrowphase(const vx_body* bodies, const Soup& soup, float dt, const std::vector<uint64_t>& keys,
                 std::vector<vx_contact>& out) {
    out.clear();
    vx_contact buf[kMaxContactsPerPair];
    for (uint64_t k : keys) {
        uint32_t a = uint32_t(k >> 32), b = uint32_t(k);
        uint32_t i = a < b ? a : b, j = a < b ? b : a;
        int cnt = collidePair(bodies[i], bodies[j], i, j, soup, dt, buf, kMaxContactsPerPair);
        // Ties (same pair, same point) keep emission order.
        std::stable_sort(buf, buf + cnt, canonicalLess);
        out.insert(out.end(), buf, buf + cnt);
    }
}

// world.cpp:4111-4128
void wakePass(vx_body* bodies, const std::vector<vx_contact>& cs) {
    for (const vx_contact& c : cs) {
        vx_body& a = bodies[c.a];
        vx_body& b = bodies[c.b];
        if (isSensor(a) || isSensor(b)) continue;
        bool impact = c.vn0 < -0.1f;
        auto moving = [](const vx_body& x) {
            return lengthSq(v3(x.velocity)) + lengthSq(v3(x.angularVelocity)) > 1e-3f;
        };
        if (asleepFully(a) && !isStatic(b) && !asleepFully(b) && (impact || moving(b))) { a.asleep = 0; a.sleepTimer = 0.0f; }
        if (asleepFully(b) && !isStatic(a) && !asleepFully(a) && (impact || moving(a))) { b.asleep = 0; b.sleepTimer = 0.0f; }
    }
}

inline void copyImpulses(vx_contact& c, const vx_contact& p) {
    c.normalImpulse = p.normalImpulse; c.tangentImpulse1 = p.tangentImpulse1; c.tangentImpulse2 = p.tangentImpulse2;
    c.rollingImpulse1 = p.rollingImpulse1; c.rollingImpulse2 = p.rollingImpulse2; c.spinningImpulse = p.spinningImpulse;
}

// world.cpp:4133-4184; prev must be sorted by (a<<32|b).
void warmStartMatch(vx_contact* cur, uint32_t m, const vx_contact* prev, uint32_t mp) {
    if (mp == 0) return;
    auto keyOf = [](const vx_contact& c) { return (uint64_t)c.a << 32 | c.b; };
    for (uint32_t k = 0; k < m; ++k) {
        vx_contact& c = cur[k];
        uint64_t key = keyOf(c);
        uint32_t lo = 0, hi = mp;
        while (lo < hi) { uint32_t mid = (lo + hi) / 2; if (keyOf(prev[mid]) < key) lo = mid + 1; else hi = mid; }
        if (c.featureKey != 0) {
            bool matched = false;
            for (uint32_t t = lo; t < mp && keyOf(prev[t]) == key; ++t)
                if (prev[t].featureKey == c.featureKey) {
                    copyImpulses(c, prev[t]);
                    if (c.vn0 < -1.0f && prev[t].vn0 < c.vn0) c.vn0 = prev[t].vn0;
                    matched = true;
                    break;
                }
            if (matched) continue;
        }
        if (c.vn0 < -1.0f)
            for (uint32_t t = lo; t < mp && keyOf(prev[t]) == key; ++t)
                if (prev[t].vn0 < c.vn0) c.vn0 = prev[t].vn0;
        float best = 2.5e-3f;
        for (uint32_t t = lo; t < mp && keyOf(prev[t]) == key; ++t) {
            float d = lengthSq(v3(prev[t].point) - v3(c.point));
            if (d < best) { best = d; copyImpulses(c, prev[t]); }
        }
    }
}

// backend_cuda.cu:603-630: greedy first-free bit, spill counters above 64.
uint32_t greedyColor(const vx_body* bodies, uint32_t n, const vx_contact* cs, uint32_t m, uint32_t* colors) {
    std::vector<uint64_t> mask(n, 0);
    std::vector<int> next(n, 64);
    int numColors = 0;
    for (uint32_t k = 0; k < m; ++k) {
        const vx_contact& c = cs[k];
        bool sensor = isSensor(bodies[c.a]) || isSensor(bodies[c.b]);
        bool sa = sensor || !isDynamic(bodies[c.a]);
        bool sb = sensor || !isDynamic(bodies[c.b]);
        uint64_t used = (sa ? 0 : mask[c.a]) | (sb ? 0 : mask[c.b]);
        int color;
        if (~used) {
            color = 0;
            while (used & (1ull << color)) ++color;
        } else {
            int na = sa ? 64 : next[c.a], nb = sb ? 64 : next[c.b];
            color = na > nb ? na : nb;
            if (!sa) next[c.a] = color + 1;
            if (!sb) next[c.b] = color + 1;
        }
        if (color < 64) {
            if (!sa) mask[c.a] |= 1ull << color;
            if (!sb) mask[c.b] |= 1ull << color;
        }
        colors[k] = (uint32_t)color;
        if (color + 1 > numColors) numColors = color + 1;
    }
    return (uint32_t)numColors;
}

// backend_cuda.cu:634-640: stable counting sort by color.
void colorOrder(const uint32_t* colors, uint32_t m, uint32_t K, std::vector<uint32_t>& order,
                std::vector<uint32_t>& start) {
    start.assign(K + 1, 0);
    for (uint32_t k = 0; k < m; ++k) ++start[colors[k] + 1];
    for (uint32_t c = 0; c < K; ++c) start[c + 1] += start[c];
    std::vector<uint32_t> fill(start.begin(), start.end());
    order.resize(m);
    for (uint32_t k = 0; k < m; ++k) order[fill[colors[k]]++] = k;
}

void solveSweeps(vx_body* bodies, vx_contact* cs, uint32_t m, float h, bool warm, int sweeps,
                 const vx_solver_options& opt) {
    if (warm)
        for (uint32_t k = 0; k < m; ++k) {
            vx_contact& c = cs[k];
            if (!isSensor(bodies[c.a]) && !isSensor(bodies[c.b])) warmStart(bodies[c.a], bodies[c.b], c);
        }
    for (int it = 0; it < sweeps; ++it)
        for (uint32_t k = 0; k < m; ++k) {
            vx_contact& c = cs[k];
            if (!isSensor(bodies[c.a]) && !isSensor(bodies[c.b])) solveContact(bodies[c.a], bodies[c.b], c, h, opt);
        }
}

// world.cpp:4364-4397
void positionCorrection(vx_body* bodies, const vx_contact* cs, uint32_t m, float h, const vx_solver_options& opt) {
    for (int it = 0; it < opt.positionIterations; ++it)
        for (uint32_t k = 0; k < m; ++k) {
            const vx_contact& c = cs[k];
            vx_body& a = bodies[c.a];
            vx_body& b = bodies[c.b];
            if (isSensor(a) || isSensor(b)) continue;
            float ims = solverInvMass(a) + solverInvMass(b);
            if (ims <= 0.0f) continue;
            float g = liveGap(a, b, c);
            if (g >= -1e-3f) continue;
            float resolve = 0.6f;
            if (opt.enableStackStabilization) {
                V3 rv = v3(a.velocity) - v3(b.velocity);
                if (lengthSq(rv) < 1e-4f) {
                    const float stiff = opt.stackStiffness / 1000.0f;
                    const float damp = 1.0f + opt.stackDamping * h;
                    resolve += 0.2f * stiff / damp;
                }
            }
            float push = -resolve * (g + 1e-3f) / ims;
            if (isDynamic(a) && !asleepFully(a)) a.position = raw(v3(a.position) + v3(c.normal) * (push * solverInvMass(a)));
            if (isDynamic(b) && !asleepFully(b)) b.position = raw(v3(b.position) - v3(c.normal) * (push * solverInvMass(b)));
        }
}

struct Gap { float gap; V3 normal; };

// world.cpp:2916-2958
Gap gapAt(const vx_body& a0, V3 pa, Q4 qa, const vx_body& b0, V3 pb, Q4 qb, const Soup& soup, float search) {
    vx_body ta = a0; ta.position = raw(pa); ta.orientation = raw(qa);
    vx_body tb = b0; tb.position = raw(pb); tb.orientation = raw(qb);
    if (ta.shape == VX_SHAPE_COMPOUND) {
        Gap best{1e30f, {0, 1, 0}};
        for (uint32_t i = 0; i < ta.compoundCount; ++i) {
            vx_body ch = childBody(ta, soup.children[ta.compoundFirst + i]);
            Gap g = gapAt(ch, v3(ch.position), q4(ch.orientation), tb, v3(tb.position), q4(tb.orientation), soup, search);
            if (g.gap < best.gap) best = g;
        }
        return best;
    }
    if (tb.shape == VX_SHAPE_COMPOUND) {
        Gap best{1e30f, {0, 1, 0}};
        for (uint32_t i = 0; i < tb.compoundCount; ++i) {
            vx_body ch = childBody(tb, soup.children[tb.compoundFirst + i]);
            Gap g = gapAt(ta, v3(ta.position), q4(ta.orientation), ch, v3(ch.position), q4(ch.orientation), soup, search);
            if (g.gap < best.gap) best = g;
        }
        return best;
    }
    if (tb.shape == VX_SHAPE_PLANE) {
        Convex c = makeConvex(ta, soup);
        V3 pn = v3(tb.planeNormal);
        V3 deep = c.support(-pn);
        return {dot(pn, deep) - tb.planeOffset - c.radius, pn};
    }
    if (tb.shape == VX_SHAPE_MESH) {
        Gap best{1e30f, {0, 1, 0}};
        meshGapProbe(ta, tb, soup, search, best.gap, best.normal);
        return best;
    }
    GjkResult r = gjkDistance(makeConvex(ta, soup), makeConvex(tb, soup));
    return {r.distance, r.normal};
}

Q4 orientationAt(const vx_body& body, Q4 start, V3 L, float elapsed) {
    vx_body s = body;
    s.orientation = raw(start);
    s.angularVelocity = raw(invInertiaMul(s, L));
    advanceOrientation(s, elapsed);
    return q4(s.orientation);
}

// world.cpp:4232-4357 (conservative advancement net; MotionQuality::High replay
// is not part of the device step).  Returns the number of rescued pairs.
uint32_t ccdNet(vx_body* bodies, const vx_vec3* prevPos, const vx_quat* prevQ, const vx_contact* cs, uint32_t m,
                const Soup& soup, float dt) {
    std::vector<uint64_t> keys;
    keys.reserve(m);
    for (uint32_t k = 0; k < m; ++k) keys.push_back((uint64_t)cs[k].a << 32 | cs[k].b);
    std::sort(keys.begin(), keys.end());
    keys.erase(std::unique(keys.begin(), keys.end()), keys.end());
    uint32_t rescued = 0;
    for (uint64_t key : keys) {
        uint32_t i = uint32_t(key >> 32), j = uint32_t(key);
        if (!isDynamic(bodies[i]) && isDynamic(bodies[j])) { uint32_t t = i; i = j; j = t; }
        vx_body& a = bodies[i];
        vx_body& b = bodies[j];
        if (isSensor(a) || isSensor(b)) continue;
        if (!isDynamic(a)) continue;
        if (!a.ccdTuning.enableContinuous || a.ccdTuning.quality == VX_QUALITY_LOW || isLocked(a) ||
            maxPointSpeed(a) < a.ccdTuning.minVelocityForCCD)
            continue;
        float search = maxPointSpeed(a) * dt + a.radius + length(v3(a.halfExtents)) + a.capsuleHalfHeight + 0.1f;
        Gap end = gapAt(a, v3(a.position), q4(a.orientation), b, v3(b.position), q4(b.orientation), soup, search);
        if (end.gap >= -8e-3f) continue;
        V3 pa0 = v3(prevPos[i]), da = v3(a.position) - pa0;
        Q4 qa0 = q4(prevQ[i]);
        V3 pb0 = isStatic(b) ? v3(b.position) : v3(prevPos[j]);
        V3 db = v3(b.position) - pb0;
        Q4 qb0 = isStatic(b) ? q4(b.orientation) : q4(prevQ[j]);
        float bound = maxPointSpeed(a) * dt + (isStatic(b) ? 0.0f : maxPointSpeed(b) * dt);
        if (bound < 1e-9f) continue;
        V3 La = angularMomentum(a);
        V3 Lb = isDynamic(b) ? angularMomentum(b) : V3{};
        float t = 0.0f;
        V3 n = end.normal;
        for (int it = 0; it < 24 && t < 1.0f; ++it) {
            Q4 qa = orientationAt(a, qa0, La, dt * t);
            Q4 qb = isStatic(b) ? qb0 : orientationAt(b, qb0, Lb, dt * t);
            Gap g = gapAt(a, pa0 + da * t, qa, b, pb0 + db * t, qb, soup, search);
            n = g.normal;
            if (g.gap < 1e-3f) break;
            t = fmn(1.0f, t + g.gap / bound);
        }
        if (t >= 1.0f) continue;
        ++rescued;
        a.position = raw(pa0 + da * t);
        a.orientation = raw(qa0);
        a.angularVelocity = raw(invInertiaMul(a, La));
        advanceOrientation(a, dt * t);
        if (isDynamic(b)) {
            b.position = raw(pb0 + db * t);
            b.orientation = raw(qb0);
            b.angularVelocity = raw(invInertiaMul(b, Lb));
            advanceOrientation(b, dt * t);
        }
        float vn = dot(v3(a.velocity) - v3(b.velocity), n);
        float ims = solverInvMass(a) + solverInvMass(b);
        if (vn < 0.0f && ims > 0.0f) {
            float rest = vn < -1.0f ? combine(a.restitution, b.restitution, a.restitutionCombine, b.restitutionCombine) : 0.0f;
            float imp = -(1.0f + rest) * vn / ims;
            a.velocity = raw(v3(a.velocity) + n * (imp * solverInvMass(a)));
            if (isDynamic(b)) b.velocity = raw(v3(b.velocity) - n * (imp * solverInvMass(b)));
        }
        float remaining = dt * (1.0f - t);
        advanceTransform(a, remaining);
        if (isDynamic(b)) advanceTransform(b, remaining);
    }
    return rescued;
}

}  // namespace

// ------------------------------------------------------------------ C API --
extern "C" {

typedef struct {
    float gravity[3];
    int32_t substeps;
    vx_solver_options options;
} vxo_config;

typedef struct {
    uint32_t pairs, contacts, colors, ccdRescued;
} vxo_stats;

void vxo_integrate(vx_body* b, uint32_t n, const float g[3], float dt) {
    for (uint32_t i = 0; i < n; ++i) integrateBody(b[i], V3{g[0], g[1], g[2]}, dt);
}
// world.cpp:4219-4222
void vxo_advance_transforms(vx_body* b, uint32_t n, float dt) {
    for (uint32_t i = 0; i < n; ++i)
        if (!isStatic(b[i]) && !asleepFully(b[i])) advanceTransform(b[i], dt);
}
void vxo_body_aabbs(const vx_body* b, uint32_t n, const vx_soup* s, float dt, float* lohi) {
    Soup soup = soupOf(s);
    for (uint32_t i = 0; i < n; ++i) {
        V3 lo, hi;
        aabbOf(b[i], soup, dt, lo, hi);
        float* o = lohi + 6 * i;
        o[0] = lo.x; o[1] = lo.y; o[2] = lo.z; o[3] = hi.x; o[4] = hi.y; o[5] = hi.z;
    }
}
uint32_t vxo_broadphase(const vx_body* b, uint32_t n, const vx_soup* s, float dt, uint64_t* out, uint32_t cap) {
    std::vector<uint64_t> p = broadphase(b, n, soupOf(s), dt);
    for (uint32_t k = 0; k < p.size() && k < cap; ++k) out[k] = p[k];
    return (uint32_t)p.size();
}
int vxo_collide_pair(const vx_body* b, const vx_soup* s, uint32_t i, uint32_t j, float dt, vx_contact* out) {
    return collidePair(b[i], b[j], i, j, soupOf(s), dt, out, kMaxContactsPerPair);
}
// The reference CPU backend's sphere-sphere contact (collideSpherePair,
// solver.cpp:26-52, taken by CpuBackend::findContacts, solver.cpp:478-483):
// returns -1 where it defers to the generic collidePair (coincident centres),
// 0 when the pair is beyond its reach ((|va| + |vb|) dt + 1e-3, with a 1e-5
// relative band), else the count emit() leaves (0 or 1).  Its reach omits the
// two speculative distances emit() adds, so separated pairs between the two
// reaches get no contact here but one from the generic path.
int vxo_sphere_pair_cpu(const vx_body* bodies, uint32_t i, uint32_t j, float dt, vx_contact* out) {
    const vx_body& a = bodies[i];
    const vx_body& b = bodies[j];
    const V3 delta = v3(a.position) - v3(b.position);
    const float distanceSq = dot(delta, delta);
    if (distanceSq <= 1e-10f) return -1;
    const float travel = (maxPointSpeed(a) + maxPointSpeed(b)) * dt + 1e-3f;
    const float maximumDistance = a.radius + b.radius + travel;
    const float safeDistance = maximumDistance + 1e-5f * fmx(1.0f, maximumDistance);
    if (distanceSq > safeDistance * safeDistance) return 0;
    const float distance = sqrtf(distanceSq);
    const V3 normal = delta * (1.0f / distance);
    const float gap = distance - a.radius - b.radius;
    const V3 pointA = v3(a.position) - normal * a.radius;
    const V3 pointB = v3(b.position) + normal * b.radius;
    int count = 0;
    emit(a, b, i, j, normal, (pointA + pointB) * 0.5f, gap, dt, out, kMaxContactsPerPair, count);
    return count;
}
uint32_t vxo_find_contacts(const vx_body* b, uint32_t n, const vx_soup* s, float dt, vx_contact* out, uint32_t cap) {
    Soup soup = soupOf(s);
    std::vector<vx_contact> cs;
    narrowphase(b, soup, dt, broadphase(b, n, soup, dt), cs);
    for (uint32_t k = 0; k < cs.size() && k < cap; ++k) out[k] = cs[k];
    return (uint32_t)cs.size();
}
// The host libstdc++'s std::sort of (key, value) pairs by key (the sort
// world.cpp:4518-4521 orders the warm-start cache with).
void vxo_std_sort(uint64_t* keys, uint32_t* vals, uint32_t n) {
    struct E { uint64_t k; uint32_t v; };
    std::vector<E> e(n);
    for (uint32_t i = 0; i < n; ++i) e[i] = {keys[i], vals[i]};
    std::sort(e.begin(), e.end(), [](const E& x, const E& y) { return x.k < y.k; });
    for (uint32_t i = 0; i < n; ++i) { keys[i] = e[i].k; vals[i] = e[i].v; }
}
void vxo_warm_start_match(vx_contact* cur, uint32_t m, const vx_contact* prev, uint32_t mp) {
    warmStartMatch(cur, m, prev, mp);
}
uint32_t vxo_color(const vx_body* b, uint32_t n, const vx_contact* cs, uint32_t m, uint32_t* colors) {
    return greedyColor(b, n, cs, m, colors);
}
// Colored replay: contacts must already be in color order.
void vxo_solve(vx_body* b, vx_contact* cs, uint32_t m, float h, int warm, int sweeps, const vx_solver_options* o) {
    solveSweeps(b, cs, m, h, warm != 0, sweeps, *o);
}
void vxo_position_correction(vx_body* b, const vx_contact* cs, uint32_t m, float h, const vx_solver_options* o) {
    positionCorrection(b, cs, m, h, *o);
}
uint32_t vxo_ccd(vx_body* b, const vx_vec3* prevPos, const vx_quat* prevQ, const vx_contact* cs, uint32_t m,
                 const vx_soup* s, float dt) {
    return ccdNet(b, prevPos, prevQ, cs, m, soupOf(s), dt);
}
float vxo_gap_at(const vx_body* b, uint32_t i, uint32_t j, const vx_soup* s, float search, float* normal) {
    Gap g = gapAt(b[i], v3(b[i].position), q4(b[i].orientation), b[j], v3(b[j].position), q4(b[j].orientation),
                  soupOf(s), search);
    if (normal) { normal[0] = g.normal.x; normal[1] = g.normal.y; normal[2] = g.normal.z; }
    return g.gap;
}

// One full device-step replay.  `prev`/`prevCount` hold the previous step's
// contacts in the reference's cache order (they are replaced on return); `outContacts`
// receives this step's contacts in color order (as the device leaves them).
int vxo_world_step(vx_body* bodies, uint32_t n, const vx_soup* s, const vxo_config* cfg, float dt,
                   vx_contact* prev, uint32_t* prevCount, uint32_t prevCap,
                   vx_contact* outContacts, uint32_t outCap, vxo_stats* stats) {
    Soup soup = soupOf(s);
    const vx_solver_options& opt = cfg->options;
    const int S = cfg->substeps > 0 ? cfg->substeps : 1;
    const float h = dt / S;
    const V3 g{cfg->gravity[0], cfg->gravity[1], cfg->gravity[2]};
    for (uint32_t i = 0; i < n; ++i) bodies[i].orientation = raw(qnormalize(q4(bodies[i].orientation)));
    for (uint32_t i = 0; i < n; ++i) integrateBody(bodies[i], g, h);
    std::vector<uint64_t> keys = broadphase(bodies, n, soup, dt);
    std::vector<vx_contact> cs;
    narrowphase(bodies, soup, dt, keys, cs);
    uint32_t m = (uint32_t)cs.size();
    wakePass(bodies, cs);
    warmStartMatch(cs.data(), m, prev, *prevCount);
    std::vector<vx_vec3> pp(n);
    std::vector<vx_quat> pq(n);
    for (uint32_t i = 0; i < n; ++i) { pp[i] = bodies[i].position; pq[i] = bodies[i].orientation; }
    std::vector<uint32_t> colors(m), order, start;
    uint32_t K = greedyColor(bodies, n, cs.data(), m, colors.data());
    colorOrder(colors.data(), m, K, order, start);
    std::vector<vx_contact> solve(m);
    for (uint32_t k = 0; k < m; ++k) solve[k] = cs[order[k]];
    const int sweeps = 2 * opt.velocityIterations;
    for (int st = 0; st < S; ++st) {
        if (st > 0) for (uint32_t i = 0; i < n; ++i) integrateBody(bodies[i], g, h);
        if (m) solveSweeps(bodies, solve.data(), m, h, st == 0, sweeps, opt);
        for (uint32_t i = 0; i < n; ++i)
            if (!isStatic(bodies[i]) && !asleepFully(bodies[i])) advanceTransform(bodies[i], h);
    }
    uint32_t rescued = ccdNet(bodies, pp.data(), pq.data(), solve.data(), m, soup, dt);
    positionCorrection(bodies, solve.data(), m, h, opt);
    for (uint32_t i = 0; i < n; ++i) { bodies[i].force = {0, 0, 0}; bodies[i].torque = {0, 0, 0}; }
    // Next step's warm-start source (world.cpp:4496-4521): this step's
    // contacts in solve (colour) order, ordered by std::sort on the pair key.
    // std::sort is not stable, and the order it leaves a pair's contacts in
    // decides which cached contact a tie in the next step's matching takes.
    std::vector<vx_contact> cache(solve.begin(), solve.end());
    std::sort(cache.begin(), cache.end(), [](const vx_contact& x, const vx_contact& y) {
        return ((uint64_t)x.a << 32 | x.b) < ((uint64_t)y.a << 32 | y.b);
    });
    uint32_t keep = m < prevCap ? m : prevCap;
    for (uint32_t k = 0; k < keep; ++k) prev[k] = cache[k];
    *prevCount = keep;
    for (uint32_t k = 0; k < m && k < outCap; ++k) outContacts[k] = solve[k];
    if (stats) { stats->pairs = (uint32_t)keys.size(); stats->contacts = m; stats->colors = K; stats->ccdRescued = rescued; }
    return (int)m;
}

// Contact events of one step (world.cpp:4399-4485) from the step's contacts
// as the solver left them (colour order, accumulated impulses).  prevKeys /
// *prevKeyCount: the previous step's sorted unique pair keys, replaced on
// return.  prevContacts: the previous step's contacts (cache order), the
// source of the representative of a pair whose bodies both went inert.
// Returns the event count (events beyond cap are counted, not written).
uint32_t vxo_contact_events(const vx_body* bodies, uint32_t n, const vx_contact* cs, uint32_t m, float dt,
                            uint64_t* prevKeys, uint32_t* prevKeyCount, uint32_t keyCap,
                            const vx_contact* prevContacts, uint32_t prevContactCount,
                            vx_contact_event* out, uint32_t cap) {
    auto canonicalKey = [](uint32_t a, uint32_t b) {
        const uint32_t lo = a < b ? a : b, hi = a < b ? b : a;
        return (uint64_t)lo << 32 | hi;
    };
    std::vector<std::pair<uint64_t, uint32_t>> active;
    for (uint32_t k = 0; k < m; ++k) {
        const vx_contact& c = cs[k];
        const vx_body &a = bodies[c.a], &b = bodies[c.b];
        const bool sensor = isSensor(a) || isSensor(b);
        const float lg = liveGap(a, b, c);
        const bool swept = sensor && c.vn0 < 0.0f && c.gap <= -c.vn0 * dt + 1e-3f;
        if (c.normalImpulse <= 1e-6f && lg > 1e-3f && !swept) continue;
        active.emplace_back(canonicalKey(c.a, c.b), k);
    }
    std::sort(active.begin(), active.end());
    std::vector<uint64_t> prev(prevKeys, prevKeys + *prevKeyCount), keys;
    std::vector<vx_contact_event> ev;
    auto event = [&](uint32_t lo, uint32_t hi, uint32_t type) {
        vx_contact_event e{};
        e.a = lo; e.b = hi; e.type = type;
        e.sensor = (isSensor(bodies[lo]) || isSensor(bodies[hi])) ? 1u : 0u;
        return e;
    };
    auto setRep = [](vx_contact_event& e, const vx_contact& c, uint32_t lo) {
        e.impulse = c.normalImpulse;
        e.point[0] = c.point.x; e.point[1] = c.point.y; e.point[2] = c.point.z;
        // Narrow-phase normals point from B to A; events use the canonical
        // low-to-high direction (world.cpp:4449-4451).
        e.normal[0] = c.a == lo ? -c.normal.x : c.normal.x;
        e.normal[1] = c.a == lo ? -c.normal.y : c.normal.y;
        e.normal[2] = c.a == lo ? -c.normal.z : c.normal.z;
    };
    for (size_t i = 0; i < active.size();) {
        const uint64_t key = active[i].first;
        const uint32_t lo = (uint32_t)(key >> 32), hi = (uint32_t)key;
        const bool persisted = std::binary_search(prev.begin(), prev.end(), key);
        vx_contact_event e = event(lo, hi, persisted ? 1u : 0u);
        bool set = false;
        for (; i < active.size() && active[i].first == key; ++i) {
            const vx_contact& c = cs[active[i].second];
            if (!set || c.normalImpulse > e.impulse) { set = true; setRep(e, c, lo); }
        }
        keys.push_back(key);
        ev.push_back(e);
    }
    for (uint64_t key : prev) {
        if (std::binary_search(keys.begin(), keys.end(), key)) continue;
        const uint32_t lo = (uint32_t)(key >> 32), hi = (uint32_t)key;
        if (lo >= n || hi >= n) continue;
        const bool inertLo = isStatic(bodies[lo]) || asleepFully(bodies[lo]);
        const bool inertHi = isStatic(bodies[hi]) || asleepFully(bodies[hi]);
        if (inertLo && inertHi) {
            vx_contact_event e = event(lo, hi, 1u);
            for (uint32_t k = 0; k < prevContactCount; ++k)
                if (canonicalKey(prevContacts[k].a, prevContacts[k].b) == key) { setRep(e, prevContacts[k], lo); break; }
            keys.push_back(key);
            ev.push_back(e);
            continue;
        }
        ev.push_back(event(lo, hi, 2u));
    }
    std::sort(keys.begin(), keys.end());
    keys.erase(std::unique(keys.begin(), keys.end()), keys.end());
    const uint32_t kc = (uint32_t)std::min<size_t>(keys.size(), keyCap);
    for (uint32_t k = 0; k < kc; ++k) prevKeys[k] = keys[k];
    *prevKeyCount = kc;
    for (uint32_t k = 0; k < ev.size() && k < cap; ++k) out[k] = ev[k];
    return (uint32_t)ev.size();
}

}  // extern "C"
