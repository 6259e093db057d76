// Device narrow phase: support maps and GJK/EPA (src/gjk.h), face clipping
// and feature ids (src/manifold.h), pair dispatch and speculative contact
// emission (src/narrowphase.h).  One pair per lane; the fixed-capacity
// arrays and iteration caps follow the reference because they shape results.
#pragma once
#include "vxd_math.hpp"

namespace vxd {

// ----------------------------------------------------------------- convexes --
enum ConvexKind : uint8_t { CPoint, CSegment, CBox, CTriangle, CHull, CCylinder, CCone,
                            CRoundedBox, CEllipsoid };

// Wavefront-cooperative hull support (Convex::coop): every lane of a full
// wavefront holds the same query.  Lane l scans points l, l + 64, ...; the
// lanes' best (dot, index) candidates are merged by six DPP steps (row shifts
// 1, 2, 4, 8, then row broadcasts 15 and 31) into lane 63.  The merge keeps the
// larger dot and, on equal dots, the lower index; dots that are NaN never
// win, and a NaN first dot keeps point 0.  That is exactly the point the
// one-lane scan (first point first, replaced on a strictly larger dot) picks.
constexpr uint32_t kNoPoint = 0xffffffffu;

template <int kCtrl, int kRowMask>
VXD void waveBestStep(float& t, uint32_t& i) {
    const float nt = __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(t), kCtrl, kRowMask, 0xf, false));
    const uint32_t ni = (uint32_t)__builtin_amdgcn_update_dpp((int)kNoPoint, (int)i, kCtrl, kRowMask, 0xf, false);
    if (ni != kNoPoint && (i == kNoPoint || nt > t || (nt == t && ni < i))) { t = nt; i = ni; }
}

VXD uint32_t hullSupportIndexWave(const vx_vec3* pts, uint32_t count, V3 d) {
    if (!(dot(d, v3(pts[0])) == dot(d, v3(pts[0])))) return 0;
    const uint32_t lane = __lane_id();
    float t = 0.0f;
    uint32_t i = kNoPoint;
    for (uint32_t base = 0; base < count; base += 64) {
        const uint32_t k = base + lane;
        if (k < count) {
            const float tk = dot(d, v3(pts[k]));
            if (tk == tk && (i == kNoPoint || tk > t)) { t = tk; i = k; }
        }
    }
    waveBestStep<0x111, 0xf>(t, i);  // row_shr:1
    waveBestStep<0x112, 0xf>(t, i);  // row_shr:2
    waveBestStep<0x114, 0xf>(t, i);  // row_shr:4
    waveBestStep<0x118, 0xf>(t, i);  // row_shr:8
    waveBestStep<0x142, 0xa>(t, i);  // row_bcast:15
    waveBestStep<0x143, 0xc>(t, i);  // row_bcast:31
    const uint32_t best = (uint32_t)__builtin_amdgcn_readlane((int)i, 63);
    return best < count ? best : 0u;  // lane 0 holds a valid candidate, so lane 63 does too
}

// The same support point when every lane holds one hull point in registers
// (lane k point k, count <= 64): no memory access per query.
VXD V3 hullSupportPointLanes(V3 lp, uint32_t count, V3 d) {
    auto lanePoint = [&](uint32_t k) {
        return V3{__int_as_float(__builtin_amdgcn_readlane(__float_as_int(lp.x), (int)k)),
                  __int_as_float(__builtin_amdgcn_readlane(__float_as_int(lp.y), (int)k)),
                  __int_as_float(__builtin_amdgcn_readlane(__float_as_int(lp.z), (int)k))};
    };
    const V3 p0 = lanePoint(0);
    if (!(dot(d, p0) == dot(d, p0))) return p0;
    const uint32_t lane = __lane_id();
    float t = 0.0f;
    uint32_t i = kNoPoint;
    if (lane < count) {
        const float tk = dot(d, lp);
        if (tk == tk) { t = tk; i = lane; }
    }
    waveBestStep<0x111, 0xf>(t, i);  // row_shr:1
    waveBestStep<0x112, 0xf>(t, i);  // row_shr:2
    waveBestStep<0x114, 0xf>(t, i);  // row_shr:4
    waveBestStep<0x118, 0xf>(t, i);  // row_shr:8
    waveBestStep<0x142, 0xa>(t, i);  // row_bcast:15
    waveBestStep<0x143, 0xc>(t, i);  // row_bcast:31
    const uint32_t best = (uint32_t)__builtin_amdgcn_readlane((int)i, 63);
    return lanePoint(best < count ? best : 0u);
}

struct Convex {                 // gjk.h:13-29
    ConvexKind kind = CPoint;
    V3 position;
    Q4 orientation;
    V3 halfExtents;
    V3 segA, segB, triC;
    const vx_vec3* hullPts = nullptr;
    uint32_t hullCount = 0;
    const uint32_t* hullFaces = nullptr;
    uint32_t hullFaceCount = 0;
    float radius = 0.0f;
    float geomRadius = 0.0f;
    float boundingRadius = 0.0f;
    uint8_t quality = VX_GEOM_NORMAL;
    bool coop = false;  // hull support by the whole wavefront (all lanes, same query)
    bool lanePts = false;  // coop hull of at most 64 points held one per lane in lanePt
    V3 lanePt;

    // Turns on wavefront-cooperative support; a hull of at most 64 points is
    // then loaded once, one point per lane.  Called by every lane.
    VXD void makeCoop() {
        coop = true;
        if (kind == CHull && hullCount <= 64) {
            const uint32_t lane = __lane_id();
            lanePt = v3(hullPts[lane < hullCount ? lane : 0]);
            lanePts = true;
        }
    }

    // gjk.h:30-148
    VXD V3 supportOffset(V3 dir) const {
        switch (kind) {
        case CPoint: return {};
        case CSegment: {
            V3 a = segA - position, b = segB - position;
            return dot(dir, a) > dot(dir, b) ? a : b;
        }
        case CTriangle: {
            V3 best = segA - position, b = segB - position, c = triC - position;
            if (dot(dir, b) > dot(dir, best)) best = b;
            if (dot(dir, c) > dot(dir, best)) best = c;
            return best;
        }
        case CBox: {
            V3 d = rotateInv(orientation, dir);
            V3 l{d.x >= 0 ? halfExtents.x : -halfExtents.x, d.y >= 0 ? halfExtents.y : -halfExtents.y,
                 d.z >= 0 ? halfExtents.z : -halfExtents.z};
            return rotate(orientation, l);
        }
        case CHull: {
            V3 d = rotateInv(orientation, dir);
            if (lanePts) return rotate(orientation, hullSupportPointLanes(lanePt, hullCount, d));
            if (coop) return rotate(orientation, v3(hullPts[hullSupportIndexWave(hullPts, hullCount, d)]));
            V3 bestP{};
            float bestDot = 0.0f;
            // Eight points per batch, loaded together (one memory round trip
            // instead of eight; the last batch masked to the count), compared
            // in index order: the first point starts the scan and the first
            // maximum wins, as in the one-at-a-time scan.
            for (uint32_t i = 0; i < hullCount; i += 8) {
                V3 p[8];
#pragma unroll
                for (int q = 0; q < 8; ++q) p[q] = i + q < hullCount ? v3(hullPts[i + q]) : V3{};
#pragma unroll
                for (int q = 0; q < 8; ++q) {
                    if (i + q >= hullCount) break;
                    const float t = dot(d, p[q]);
                    if ((i == 0 && q == 0) || t > bestDot) { bestDot = t; bestP = p[q]; }
                }
            }
            return rotate(orientation, bestP);
        }
        case CCylinder: {
            V3 d = rotateInv(orientation, dir);
            float radial = sqrtf(d.x * d.x + d.z * d.z);
            V3 l{radial > 1e-9f ? geomRadius * d.x / radial : 0.0f,
                 d.y >= 0.0f ? halfExtents.y : -halfExtents.y,
                 radial > 1e-9f ? geomRadius * d.z / radial : 0.0f};
            return rotate(orientation, l);
        }
        case CCone: {
            V3 d = rotateInv(orientation, dir);
            float radial = sqrtf(d.x * d.x + d.z * d.z);
            float hh = halfExtents.y;
            V3 base{radial > 1e-9f ? geomRadius * d.x / radial : 0.0f, -0.5f * hh,
                    radial > 1e-9f ? geomRadius * d.z / radial : 0.0f};
            V3 tip{0, 1.5f * hh, 0};
            return rotate(orientation, dot(d, tip) > dot(d, base) ? tip : base);
        }
        case CRoundedBox: {
            V3 d = rotateInv(orientation, dir);
            V3 core{d.x >= 0 ? halfExtents.x : -halfExtents.x, d.y >= 0 ? halfExtents.y : -halfExtents.y,
                    d.z >= 0 ? halfExtents.z : -halfExtents.z};
            V3 n = normalize(d);
            return rotate(orientation, core + n * radius);
        }
        case CEllipsoid: {
            V3 d = rotateInv(orientation, dir);
            V3 s{halfExtents.x * halfExtents.x * d.x, halfExtents.y * halfExtents.y * d.y,
                 halfExtents.z * halfExtents.z * d.z};
            float len = length(s);
            return rotate(orientation, len > 1e-9f ? s * (1.0f / len) : V3{});
        }
        }
        return {};
    }
    VXD V3 support(V3 dir) const { return position + supportOffset(dir); }
};

struct Soup {                     // backend.h:25-34
    const vx_vec3* vertices = nullptr;
    const uint32_t* indices = nullptr;
    const vx_mesh* meshes = nullptr;
    const vx_bvh_node* nodes = nullptr;
    const uint32_t* triRefs = nullptr;
    const vx_vec3* hullPoints = nullptr;
    const uint32_t* hullFaceIndices = nullptr;
    const vx_compound_child* children = nullptr;
};

// gjk.h:155-226
VXD Convex makeConvex(const vx_body& b, const Soup& soup) {
    Convex c;
    c.position = v3(b.position);
    c.orientation = q4(b.orientation);
    c.quality = b.geometryQuality;
    switch (b.shape) {
    case VX_SHAPE_SPHERE: c.kind = CPoint; c.radius = b.radius; break;
    case VX_SHAPE_CAPSULE: {
        c.kind = CSegment;
        V3 axis = rotate(c.orientation, V3{0, b.capsuleHalfHeight, 0});
        c.segA = c.position + axis;
        c.segB = c.position - axis;
        c.radius = b.radius;
        c.boundingRadius = b.capsuleHalfHeight;
        break;
    }
    case VX_SHAPE_BOX:
        c.kind = CBox; c.halfExtents = v3(b.halfExtents);
        c.boundingRadius = length(c.halfExtents);
        break;
    case VX_SHAPE_HULL:
        c.kind = CHull;
        c.hullPts = soup.hullPoints + b.hullFirst;
        c.hullCount = b.hullCount;
        c.hullFaces = b.hullFaceCount ? soup.hullFaceIndices + b.hullFaceFirst : nullptr;
        c.hullFaceCount = b.hullFaceCount;
        c.boundingRadius = b.radius;
        break;
    case VX_SHAPE_CYLINDER:
        c.kind = CCylinder; c.halfExtents.y = b.capsuleHalfHeight; c.geomRadius = b.radius;
        c.boundingRadius = sqrtf(b.capsuleHalfHeight * b.capsuleHalfHeight + b.radius * b.radius);
        break;
    case VX_SHAPE_CONE:
        c.kind = CCone; c.halfExtents.y = b.capsuleHalfHeight; c.geomRadius = b.radius;
        c.boundingRadius = fmx(1.5f * b.capsuleHalfHeight,
                               sqrtf(b.radius * b.radius + 0.25f * b.capsuleHalfHeight * b.capsuleHalfHeight));
        break;
    case VX_SHAPE_ROUNDED_BOX:
        c.kind = CRoundedBox; c.halfExtents = v3(b.halfExtents); c.radius = b.radius;
        c.boundingRadius = length(c.halfExtents) + b.radius;
        break;
    case VX_SHAPE_ELLIPSOID:
        c.kind = CEllipsoid; c.halfExtents = v3(b.halfExtents);
        c.boundingRadius = fmx(b.halfExtents.x, fmx(b.halfExtents.y, b.halfExtents.z));
        break;
    default: c.kind = CPoint; break;
    }
    return c;
}

// gjk.h:228-239
VXD Convex makeTriangle(V3 a, V3 b, V3 c) {
    Convex t;
    t.kind = CTriangle;
    t.position = a + ((b - a) + (c - a)) * (1.0f / 3.0f);
    t.segA = a; t.segB = b; t.triC = c;
    t.boundingRadius = fmx(length(a - t.position), fmx(length(b - t.position), length(c - t.position)));
    return t;
}

struct GjkResult { float distance = 0; V3 normal, pointA, pointB; bool overlapping = false; };

// gjk.h:251-297
VXD uint8_t pairQuality(const Convex& A, const Convex& B) { return A.quality >= B.quality ? A.quality : B.quality; }
VXD float convexScale(const Convex& c) {
    if (c.boundingRadius > 0.0f) return c.boundingRadius + c.radius;
    if (c.kind == CBox) return length(c.halfExtents) + c.radius;
    if (c.kind == CSegment) return 0.5f * length(c.segA - c.segB) + c.radius;
    if (c.kind == CTriangle)
        return fmx(length(c.segA - c.position), fmx(length(c.segB - c.position), length(c.triC - c.position)));
    return c.radius;
}
VXD float pairScale(const Convex& A, const Convex& B) { return fmx(1e-8f, convexScale(A) + convexScale(B)); }
VXD float relTol(const Convex& A, const Convex& B) {
    uint8_t q = pairQuality(A, B);
    return q == VX_GEOM_NORMAL ? 1e-7f : q == VX_GEOM_ROBUST ? 3e-7f : 1e-6f;
}
VXD float dupDistSq(const Convex& A, const Convex& B) {
    if (pairQuality(A, B) == VX_GEOM_NORMAL) return 1e-12f;
    float e = pairScale(A, B) * relTol(A, B);
    return fmx(1e-30f, e * e);
}
VXD float witnessResidual(const Convex& A, const Convex& B, const GjkResult& r) {
    if (r.overlapping || lengthSq(r.normal) < 0.25f) return 1e30f;
    float aPlane = dot(-r.normal, A.support(-r.normal)) - A.radius;
    float bPlane = dot(r.normal, B.support(r.normal)) + B.radius;
    return fabsf(dot(-r.normal, r.pointA) - aPlane) + fabsf(dot(r.normal, r.pointB) - bPlane);
}

struct SV { V3 p, a, b; };  // Minkowski point + witnesses (gjk.h:287-291)
VXD SV supportVertex(const Convex& A, const Convex& B, V3 dir) {
    SV v;
    V3 oa = A.supportOffset(dir), ob = B.supportOffset(-dir);
    v.a = A.position + oa;
    v.b = B.position + ob;
    v.p = (A.position - B.position) + (oa - ob);
    return v;
}
struct Simplex { SV v[4]; float bary[4] = {0, 0, 0, 0}; int count = 0; };

// Simplex moves are written as per-component value selects, never as copies
// between slots chosen by a branch: the compiler turns branch-chosen copies
// into copies through a pointer chosen at run time, which keeps the simplex
// in private (scratch) memory instead of registers.
VXD V3 selV3(bool c, V3 x, V3 y) { return {c ? x.x : y.x, c ? x.y : y.y, c ? x.z : y.z}; }
VXD SV selSV(bool c, const SV& x, const SV& y) { return {selV3(c, x.p, y.p), selV3(c, x.a, y.a), selV3(c, x.b, y.b)}; }
VXD Simplex selSimplex(bool c, const Simplex& x, const Simplex& y) {
    Simplex r;
#pragma unroll
    for (int i = 0; i < 4; ++i) {
        r.v[i] = selSV(c, x.v[i], y.v[i]);
        r.bary[i] = c ? x.bary[i] : y.bary[i];
    }
    r.count = c ? x.count : y.count;
    return r;
}

// gjk.h:317-325
VXD void solveSegment(Simplex& s) {
    V3 a = s.v[0].p, b = s.v[1].p, ab = b - a;
    float t = -dot(a, ab);
    const float denom = dot(ab, ab);
    const bool atA = t <= 0.0f, atB = !atA && t >= denom;
    if (!atA && !atB) t /= denom;
    s.v[0] = selSV(atB, s.v[1], s.v[0]);
    s.count = atA || atB ? 1 : 2;
    s.bary[0] = atA || atB ? 1.0f : 1.0f - t;
    s.bary[1] = atA || atB ? s.bary[1] : t;
}

// gjk.h:327-376 (Ericson's closest point on triangle with a degenerate guard)
VXD void solveTriangle(Simplex& s) {
    // Region codes: which vertices the reduced simplex keeps (slot 0 from
    // src0, slot 1 from src1), its size and weights.
    int src0 = 0, src1 = 1, count = 3;
    float w0 = 1.0f, w1 = 0.0f, w2 = 0.0f;
    V3 a = s.v[0].p, b = s.v[1].p, c = s.v[2].p;
    V3 ab = b - a, ac = c - a, ap = -a;
    float d1 = dot(ab, ap), d2 = dot(ac, ap);
    bool done = false;
    if (d1 <= 0.0f && d2 <= 0.0f) { count = 1; w0 = 1.0f; done = true; }
    V3 bp = -b;
    float d3 = dot(ab, bp), d4 = dot(ac, bp);
    if (!done && d3 >= 0.0f && d4 <= d3) { src0 = 1; count = 1; w0 = 1.0f; done = true; }
    float vc = d1 * d4 - d3 * d2;
    if (!done && vc <= 0.0f && d1 >= 0.0f && d3 <= 0.0f) {
        float t = d1 / (d1 - d3);
        count = 2; w0 = 1.0f - t; w1 = t; done = true;
    }
    V3 cp = -c;
    float d5 = dot(ab, cp), d6 = dot(ac, cp);
    if (!done && d6 >= 0.0f && d5 <= d6) { src0 = 2; count = 1; w0 = 1.0f; done = true; }
    float vb = d5 * d2 - d1 * d6;
    if (!done && vb <= 0.0f && d2 >= 0.0f && d6 <= 0.0f) {
        float t = d2 / (d2 - d6);
        src1 = 2; count = 2; w0 = 1.0f - t; w1 = t; done = true;
    }
    float va = d3 * d6 - d5 * d4;
    if (!done && va <= 0.0f && (d4 - d3) >= 0.0f && (d5 - d6) >= 0.0f) {
        float t = (d4 - d3) / ((d4 - d3) + (d5 - d6));
        src0 = 1; src1 = 2; count = 2; w0 = 1.0f - t; w1 = t; done = true;
    }
    if (!done) {
        float sum = va + vb + vc;
        if (sum < 1e-20f) {
            // Nearest of the three vertices (first on ties).
            const float l0 = lengthSq(s.v[0].p), l1 = lengthSq(s.v[1].p), l2 = lengthSq(s.v[2].p);
            const int best01 = l1 < l0 ? 1 : 0;
            const float lb = best01 ? l1 : l0;
            src0 = l2 < lb ? 2 : best01;
            count = 1; w0 = 1.0f;
        } else {
            float inv = 1.0f / sum;
            count = 3; w0 = va * inv; w1 = vb * inv; w2 = vc * inv;
        }
    }
    const SV v0 = s.v[0], v1 = s.v[1], v2 = s.v[2];
    s.v[0] = selSV(src0 == 2, v2, selSV(src0 == 1, v1, v0));
    s.v[1] = selSV(src1 == 2, v2, v1);
    s.count = count;
    s.bary[0] = w0;
    if (count >= 2) s.bary[1] = w1;
    if (count == 3) s.bary[2] = w2;
}

// Loops over simplex vertices are unrolled with a count guard so that every
// index is a compile-time constant and the simplex stays in registers.
VXD V3 simplexPoint(const Simplex& s) {
    V3 c{};
#pragma unroll
    for (int i = 0; i < 4; ++i) {
        if (i >= s.count) break;
        c += s.v[i].p * s.bary[i];
    }
    return c;
}

VXD void simplexPush(Simplex& s, const SV& w) {
    const int k = s.count;
    s.v[0] = selSV(k == 0, w, s.v[0]);
    s.v[1] = selSV(k == 1, w, s.v[1]);
    s.v[2] = selSV(k == 2, w, s.v[2]);
    s.v[3] = selSV(k >= 3, w, s.v[3]);
    ++s.count;
}

__device__ constexpr int kTetFaces[4][3] = {{0, 1, 2}, {0, 1, 3}, {0, 2, 3}, {1, 2, 3}};

// The face (i0, i1, i2) of the tetrahedron as a triangle simplex.
template <int i0, int i1, int i2>
VXD Simplex tetFace(const Simplex& s) {
    Simplex tri;
    tri.count = 3;
    tri.v[0] = s.v[i0]; tri.v[1] = s.v[i1]; tri.v[2] = s.v[i2];
    return tri;
}

// gjk.h:379-445 (faces visited in kTetFaces order)
VXD bool solveTetrahedron(Simplex& s, float scale) {
    float volume = dot(s.v[1].p - s.v[0].p, cross(s.v[2].p - s.v[0].p, s.v[3].p - s.v[0].p));
    float floorV = 1e-10f * scale * scale * scale;
    if (fabsf(volume) < floorV) {
        Simplex best; float bestD = 1e30f;
        auto visit = [&](Simplex tri) {
            solveTriangle(tri);
            float d = lengthSq(simplexPoint(tri));
            const bool better = d < bestD;
            bestD = better ? d : bestD;
            best = selSimplex(better, tri, best);
        };
        visit(tetFace<0, 1, 2>(s));
        visit(tetFace<0, 1, 3>(s));
        visit(tetFace<0, 2, 3>(s));
        visit(tetFace<1, 2, 3>(s));
        s = best;
        return false;
    }
    Simplex best; float bestD = 1e30f;
    bool inside = true;
    const float band = 3e-6f * scale;
    auto face = [&](const Simplex& tri, V3 opp) {
        V3 a = tri.v[0].p, b = tri.v[1].p, c = tri.v[2].p;
        V3 n = cross(b - a, c - a);
        float nl = length(n);
        if (nl < 1e-20f) return;
        float od = dot(n, -a) / nl;
        float pd = dot(n, opp - a) / nl;
        if (od * pd < 0.0f || fabsf(od) <= band) {
            inside = false;
            Simplex t = tri;
            solveTriangle(t);
            float d = lengthSq(simplexPoint(t));
            const bool better = d < bestD;
            bestD = better ? d : bestD;
            best = selSimplex(better, t, best);
        }
    };
    face(tetFace<0, 1, 2>(s), s.v[3].p);
    face(tetFace<0, 1, 3>(s), s.v[2].p);
    face(tetFace<0, 2, 3>(s), s.v[1].p);
    face(tetFace<1, 2, 3>(s), s.v[0].p);
    s = selSimplex(!inside, best, s);
    return inside;
}

// gjk.h:447-469
VXD bool originInTet(const SV& a, const SV& b, const SV& c, const SV& d, float scale) {
    V3 v[4] = {a.p, b.p, c.p, d.p};
    float volume = dot(v[1] - v[0], cross(v[2] - v[0], v[3] - v[0]));
    float floorV = 1e-10f * scale * scale * scale;
    if (fabsf(volume) < floorV) return false;
#pragma unroll
    for (int f = 0; f < 4; ++f) {
        int i0 = kTetFaces[f][0], i1 = kTetFaces[f][1], i2 = kTetFaces[f][2];
        V3 n = cross(v[i1] - v[i0], v[i2] - v[i0]);
        float os = dot(n, -v[i0]);
        float ps = dot(n, v[6 - i0 - i1 - i2] - v[i0]);
        if (os * ps <= floorV * floorV) return false;
    }
    return true;
}

// gjk.h:474-502
VXD bool expandOverlapSimplex(const Convex& A, const Convex& B, Simplex& s) {
    SV cand[18];
    int count = 0;
#pragma unroll
    for (int i = 0; i < 4; ++i) {
        if (i >= s.count) break;
        cand[count++] = s.v[i];
    }
    constexpr float dirs[14][3] = {{1, 0, 0}, {-1, 0, 0}, {0, 1, 0}, {0, -1, 0}, {0, 0, 1}, {0, 0, -1},
                                      {1, 1, 1}, {-1, 1, 1}, {1, -1, 1}, {1, 1, -1},
                                      {-1, -1, 1}, {-1, 1, -1}, {1, -1, -1}, {-1, -1, -1}};
    for (int d = 0; d < 14 && count < 18; ++d) {
        SV w = supportVertex(A, B, V3{dirs[d][0], dirs[d][1], dirs[d][2]});
        bool dup = false;
        for (int i = 0; i < count; ++i)
            if (lengthSq(w.p - cand[i].p) < dupDistSq(A, B)) dup = true;
        if (!dup) cand[count++] = w;
    }
    for (int i = 0; i < count - 3; ++i)
        for (int j = i + 1; j < count - 2; ++j)
            for (int k = j + 1; k < count - 1; ++k)
                for (int l = k + 1; l < count; ++l)
                    if (originInTet(cand[i], cand[j], cand[k], cand[l], pairScale(A, B))) {
                        s.count = 4;
                        s.v[0] = cand[i]; s.v[1] = cand[j]; s.v[2] = cand[k]; s.v[3] = cand[l];
                        return true;
                    }
    return false;
}

// Vertex indices are below 64, so faces and horizon edges store them as
// bytes (a smaller private footprint per lane; identical results).
struct EpaFace { uint8_t a, b, c, valid; V3 normal; float distance; };

// gjk.h:511-525
VXD bool makeEpaFace(const SV* vs, int ia, int ib, int ic, float minAreaSq, EpaFace& f) {
    V3 a = vs[ia].p, b = vs[ib].p, c = vs[ic].p;
    V3 n = cross(b - a, c - a);
    float n2 = lengthSq(n);
    if (n2 < minAreaSq) return false;
    n *= 1.0f / sqrtf(n2);
    float dist = dot(n, a);
    if (dist < 0.0f) { int t = ib; ib = ic; ic = t; n = -n; dist = -dist; }
    f = {(uint8_t)ia, (uint8_t)ib, (uint8_t)ic, 1, n, dist};
    return true;
}

// gjk.h:527-542
VXD void epaWitness(const SV* vs, const EpaFace& f, V3& pa, V3& pb) {
    Simplex tri; tri.count = 3;
    tri.v[0] = vs[f.a]; tri.v[1] = vs[f.b]; tri.v[2] = vs[f.c];
    solveTriangle(tri);
    pa = {}; pb = {};
    for (int i = 0; i < tri.count; ++i) { pa += tri.v[i].a * tri.bary[i]; pb += tri.v[i].b * tri.bary[i]; }
}

VXD void epaFinish(const Convex& A, const Convex& B, const SV* vs, const EpaFace& f, GjkResult& r) {
    V3 ca, cb;
    epaWitness(vs, f, ca, cb);
    r.normal = -f.normal;
    r.distance = -(f.distance + A.radius + B.radius);
    r.pointA = ca - r.normal * A.radius;
    r.pointB = cb + r.normal * B.radius;
    r.overlapping = true;
}

// gjk.h:546-657
VXD bool epaPenetration(const Convex& A, const Convex& B, const Simplex& simplex, GjkResult& r) {
    constexpr int kV = 64, kF = 128, kE = 192;
    SV vs[kV];
    EpaFace faces[kF];
    int nv = 4, nf = 0;
    for (int i = 0; i < 4; ++i) vs[i] = simplex.v[i];
    constexpr int init[4][3] = {{0, 1, 2}, {0, 3, 1}, {0, 2, 3}, {1, 3, 2}};
    float maxArea = 0.0f;
    for (int i = 0; i < 4; ++i) {
        V3 a = vs[init[i][0]].p, b = vs[init[i][1]].p, c = vs[init[i][2]].p;
        maxArea = fmx(maxArea, lengthSq(cross(b - a, c - a)));
    }
    const float scale = pairScale(A, B);
    const uint8_t q = pairQuality(A, B);
    const float areaFloor = q == VX_GEOM_NORMAL ? 1e-16f
        : fmx(maxArea * 1e-16f, 1e-20f * scale * scale * scale * scale);
    for (int i = 0; i < 4; ++i) {
        EpaFace f;
        if (makeEpaFace(vs, init[i][0], init[i][1], init[i][2], q == VX_GEOM_NORMAL ? 1e-16f : 0.0f, f))
            faces[nf++] = f;
    }
    if (nf < 4) return false;
    int bestFace = 0;
    const int maxIt = q == VX_GEOM_NORMAL ? 48 : 64;
    for (int it = 0; it < maxIt; ++it) {
        bestFace = -1;
        float bestD = 1e30f;
        for (int i = 0; i < nf; ++i)
            if (faces[i].valid && faces[i].distance < bestD) { bestD = faces[i].distance; bestFace = i; }
        if (bestFace < 0) return false;
        EpaFace closest = faces[bestFace];
        SV w = supportVertex(A, B, closest.normal);
        float sd = dot(closest.normal, w.p);
        float tol = q == VX_GEOM_NORMAL ? 1e-4f * fmx(1.0f, sd) : 1e-4f * fmx(scale, fabsf(sd));
        bool dup = false;
        for (int i = 0; i < nv; ++i)
            if (lengthSq(w.p - vs[i].p) < dupDistSq(A, B)) dup = true;
        if (sd - closest.distance <= tol || dup) { epaFinish(A, B, vs, closest, r); return true; }
        if (nv >= kV) break;
        int nvtx = nv++;
        vs[nvtx] = w;
        uint8_t ea[kE], eb[kE];
        int ne = 0;
        for (int i = 0; i < nf; ++i) {
            if (!faces[i].valid) continue;
            const EpaFace f = faces[i];
            if (dot(f.normal, w.p - vs[f.a].p) <= tol) continue;
            faces[i].valid = 0;
            uint8_t fe[3][2] = {{f.a, f.b}, {f.b, f.c}, {f.c, f.a}};
            for (int e = 0; e < 3; ++e) {
                int rev = -1;
                for (int k = 0; k < ne; ++k)
                    if (ea[k] == fe[e][1] && eb[k] == fe[e][0]) { rev = k; break; }
                if (rev >= 0) { --ne; ea[rev] = ea[ne]; eb[rev] = eb[ne]; }
                else if (ne < kE) { ea[ne] = fe[e][0]; eb[ne] = fe[e][1]; ++ne; }
            }
        }
        for (int e = 0; e < ne; ++e) {
            EpaFace f;
            if (!makeEpaFace(vs, ea[e], eb[e], nvtx, areaFloor, f)) continue;
            int slot = -1;
            for (int i = 0; i < nf; ++i) if (!faces[i].valid) { slot = i; break; }
            if (slot >= 0) faces[slot] = f;
            else if (nf < kF) faces[nf++] = f;
            else return false;
        }
    }
    if (bestFace >= 0 && faces[bestFace].valid) { epaFinish(A, B, vs, faces[bestFace], r); return true; }
    return false;
}

// gjk.h:661-749
VXD GjkResult gjkSeeded(const Convex& A, const Convex& B, V3 dir) {
    Simplex s;
    const float scale = pairScale(A, B);
    const float scaleSq = scale * scale;
    const uint8_t q = pairQuality(A, B);
    const bool robust = q != VX_GEOM_NORMAL;
    const float tiny = robust ? scaleSq * 1e-14f : 1e-12f;
    if (lengthSq(dir) < tiny) dir = A.position - B.position;
    if (lengthSq(dir) < tiny) dir = {1, 0, 0};
    s.v[0] = supportVertex(A, B, dir);
    s.count = 1;
    s.bary[0] = 1.0f;
    GjkResult r;
    const int maxIt = q == VX_GEOM_NORMAL ? 32 : q == VX_GEOM_ROBUST ? 64 : 96;
    for (int it = 0; it < maxIt; ++it) {
        V3 closest = simplexPoint(s);
        float distSq = lengthSq(closest);
        if (distSq < (robust ? scaleSq * 1e-10f : 1e-10f)) {
            if (s.count < 4) expandOverlapSimplex(A, B, s);
            r.overlapping = true;
            break;
        }
        V3 d = -closest;
        SV w = supportVertex(A, B, d);
        float progress = dot(w.p, d) - dot(closest, d);
        float pf = robust ? relTol(A, B) * scale * sqrtf(distSq) + relTol(A, B) * relTol(A, B) * scaleSq
                          : 1e-6f * sqrtf(distSq) + 1e-9f;
        if (progress < pf) break;
        bool dup = false;
#pragma unroll
        for (int i = 0; i < 4; ++i) {
            if (i >= s.count) break;
            if (lengthSq(w.p - s.v[i].p) < dupDistSq(A, B)) { dup = true; break; }
        }
        if (dup) break;
        simplexPush(s, w);
        if (s.count == 2) solveSegment(s);
        else if (s.count == 3) solveTriangle(s);
        else if (solveTetrahedron(s, scale)) { r.overlapping = true; break; }
    }
    V3 pa{}, pb{};
#pragma unroll
    for (int i = 0; i < 4; ++i) {
        if (i >= s.count) break;
        pa += s.v[i].a * s.bary[i];
        pb += s.v[i].b * s.bary[i];
    }
    if (r.overlapping) {
        if (s.count == 4 && epaPenetration(A, B, s, r)) return r;
        V3 n = A.position - B.position;
        n = lengthSq(n) > tiny ? normalize(n) : V3{0, 1, 0};
        float depth = dot(n, B.support(n) - A.support(-n));
        r.normal = n;
        r.distance = -depth - A.radius - B.radius;
        r.pointA = A.support(-n);
        r.pointB = B.support(n);
        return r;
    }
    V3 delta = pa - pb;
    float core = length(delta);
    r.normal = core > (robust ? scale * 1e-7f : 1e-8f) ? delta * (1.0f / core) : V3{0, 1, 0};
    r.distance = core - A.radius - B.radius;
    r.pointA = pa - r.normal * A.radius;
    r.pointB = pb + r.normal * B.radius;
    return r;
}

// gjk.h:751-777
VXD GjkResult gjkDistance(const Convex& A, const Convex& B) {
    GjkResult best = gjkSeeded(A, B, A.position - B.position);
    const uint8_t q = pairQuality(A, B);
    if (q == VX_GEOM_NORMAL || best.overlapping) return best;
    constexpr float dirs[14][3] = {{1, 0, 0}, {-1, 0, 0}, {0, 1, 0}, {0, -1, 0}, {0, 0, 1}, {0, 0, -1},
                                      {1, 1, 1}, {-1, 1, 1}, {1, -1, 1}, {1, 1, -1},
                                      {-1, -1, 1}, {-1, 1, -1}, {1, -1, -1}, {-1, -1, -1}};
    const int nd = q == VX_GEOM_ROBUST ? 6 : 14;
    float bestRes = witnessResidual(A, B, best);
    for (int i = 0; i < nd; ++i) {
        GjkResult cand = gjkSeeded(A, B, V3{dirs[i][0], dirs[i][1], dirs[i][2]});
        if (cand.overlapping) return cand;
        float res = witnessResidual(A, B, cand);
        float slack = pairScale(A, B) * 1e-6f;
        if (res + slack < bestRes || (res <= bestRes + slack && cand.distance < best.distance)) {
            best = cand; bestRes = res;
        }
    }
    return best;
}

// ----------------------------------------------------------- face clipping --
constexpr uint32_t kEdgeTag = 0x20000000u;   // manifold.h:38-39
constexpr uint32_t kFaceTag = 0x40000000u;
struct ClipV { V3 p; uint32_t id; };
struct FacePoly { ClipV v[64]; int count = 0; uint32_t faceId = kFaceTag; float plane = 0.0f; };
struct Clip2 { float u, v; uint32_t fa, fb; };
struct MPoint { V3 normal, pointA, pointB; float gap; uint32_t fa, fb; };

// manifold.h:53-59
VXD uint32_t edgeFeature(uint32_t a, uint32_t b) {
    uint32_t lo = a < b ? a : b, hi = a < b ? b : a;
    return kEdgeTag | (((lo * 0x1f123bb5u) ^ (hi * 0x5f356495u)) & 0x1fffffffu);
}
// manifold.h:61-73: FNV-1a over the sorted vertex ids.
VXD uint32_t faceFeature(const uint32_t* ids, int count) {
    uint32_t s[64];
    for (int i = 0; i < count; ++i) s[i] = ids[i];
    for (int i = 1; i < count; ++i) {
        uint32_t k = s[i]; int j = i - 1;
        while (j >= 0 && s[j] > k) { s[j + 1] = s[j]; --j; }
        s[j + 1] = k;
    }
    uint32_t h = 2166136261u;
    for (int i = 0; i < count; ++i) h = (h ^ s[i]) * 16777619u;
    return kFaceTag | (h & 0x1fffffffu);
}
// manifold.h:75-79
VXD void planeBasis(V3 n, V3& u, V3& v) {
    V3 ref = fabsf(n.x) < 0.9f ? V3{1, 0, 0} : V3{0, 1, 0};
    u = normalize(cross(n, ref));
    v = cross(n, u);
}
// manifold.h:81-100: insertion sort by polar angle about the centroid.
VXD void sortCcw(ClipV* vs, int count, V3 u, V3 v) {
    if (count < 2) return;
    V3 c{};
    for (int i = 0; i < count; ++i) c += vs[i].p;
    c *= 1.0f / float(count);
    for (int i = 1; i < count; ++i) {
        ClipV key = vs[i];
        float ka = atan2f(dot(key.p - c, v), dot(key.p - c, u));
        int j = i - 1;
        while (j >= 0) {
            float a = atan2f(dot(vs[j].p - c, v), dot(vs[j].p - c, u));
            if (a <= ka) break;
            vs[j + 1] = vs[j];
            --j;
        }
        vs[j + 1] = key;
    }
}
// manifold.h:102-128
VXD int boxFace(const Convex& cv, V3 dir, FacePoly& out) {
    V3 d = rotateInv(cv.orientation, dir);
    float ax = fabsf(d.x), ay = fabsf(d.y), az = fabsf(d.z);
    int axis = ax >= ay && ax >= az ? 0 : (ay >= az ? 1 : 2);
    bool pos = (axis == 0 ? d.x : axis == 1 ? d.y : d.z) >= 0.0f;
    int o0 = axis == 0 ? 1 : 0, o1 = axis == 2 ? 1 : 2;
    out.count = 0;
    float ext[3] = {cv.halfExtents.x, cv.halfExtents.y, cv.halfExtents.z};
    for (int bits = 0; bits < 4; ++bits) {
        float val[3] = {-ext[0], -ext[1], -ext[2]};
        val[axis] = pos ? ext[axis] : -ext[axis];
        val[o0] = (bits & 1) ? ext[o0] : -ext[o0];
        val[o1] = (bits & 2) ? ext[o1] : -ext[o1];
        int vid = (val[0] > 0.0f ? 1 : 0) | (val[1] > 0.0f ? 2 : 0) | (val[2] > 0.0f ? 4 : 0);
        out.v[out.count++] = {cv.position + rotate(cv.orientation, V3{val[0], val[1], val[2]}), (uint32_t)vid};
    }
    out.faceId = kFaceTag | uint32_t(axis * 2 + (pos ? 1 : 0));
    return out.count;
}
// manifold.h:130-217
VXD int hullFace(const Convex& cv, V3 dir, FacePoly& out) {
    if (!cv.hullPts || cv.hullCount == 0) return 0;
    if (cv.hullFaces && cv.hullFaceCount > 0) {
        int seed = -1;
        V3 seedN{};
        float bestAl = -1e30f;
        for (uint32_t f = 0; f < cv.hullFaceCount; ++f) {
            const uint32_t* ix = cv.hullFaces + f * 3;
            if (ix[0] >= cv.hullCount || ix[1] >= cv.hullCount || ix[2] >= cv.hullCount) return 0;
            V3 ln = cross(v3(cv.hullPts[ix[1]]) - v3(cv.hullPts[ix[0]]), v3(cv.hullPts[ix[2]]) - v3(cv.hullPts[ix[0]]));
            if (lengthSq(ln) < 1e-16f) return 0;
            V3 wn = rotate(cv.orientation, normalize(ln));
            float al = dot(dir, wn);
            if (al > bestAl) { bestAl = al; seed = (int)f; seedN = wn; }
        }
        if (seed < 0 || bestAl < 0.95f) return 0;
        const uint32_t* si = cv.hullFaces + seed * 3;
        V3 seedP = cv.position + rotate(cv.orientation, v3(cv.hullPts[si[0]]));
        float scale = 1.0f;
        for (uint32_t i = 0; i < cv.hullCount; ++i) scale = fmx(scale, length(v3(cv.hullPts[i])));
        float ptol = 1e-5f + 1e-4f * scale;
        out.count = 0;
        uint32_t ids[64];
        for (uint32_t f = 0; f < cv.hullFaceCount; ++f) {
            const uint32_t* ix = cv.hullFaces + f * 3;
            V3 ln = cross(v3(cv.hullPts[ix[1]]) - v3(cv.hullPts[ix[0]]), v3(cv.hullPts[ix[2]]) - v3(cv.hullPts[ix[0]]));
            if (lengthSq(ln) < 1e-16f) return 0;
            V3 wn = rotate(cv.orientation, normalize(ln));
            V3 p = cv.position + rotate(cv.orientation, v3(cv.hullPts[ix[0]]));
            if (dot(seedN, wn) < 1.0f - 1e-4f || fabsf(dot(seedN, p - seedP)) > ptol) continue;
            for (int k = 0; k < 3; ++k) {
                uint32_t id = ix[k];
                bool present = false;
                for (int i = 0; i < out.count; ++i) if (out.v[i].id == id) present = true;
                if (!present) {
                    if (out.count >= 64) return 0;
                    out.v[out.count] = {cv.position + rotate(cv.orientation, v3(cv.hullPts[id])), id};
                    ids[out.count++] = id;
                }
            }
        }
        if (out.count < 3) return 0;
        out.faceId = faceFeature(ids, out.count);
        return out.count;
    }
    float sup = -1e30f, mins = 1e30f;
    V3 world[64];
    int total = cv.hullCount < 64u ? (int)cv.hullCount : 64;
    for (int i = 0; i < total; ++i) {
        world[i] = cv.position + rotate(cv.orientation, v3(cv.hullPts[i]));
        float pr = dot(dir, world[i]);
        sup = fmx(sup, pr); mins = fmn(mins, pr);
    }
    float tol = 1e-5f + 1e-4f * fmx(1.0f, sup - mins);
    out.count = 0;
    uint32_t ids[64];
    for (int i = 0; i < total; ++i)
        if (dot(dir, world[i]) >= sup - tol) { out.v[out.count] = {world[i], (uint32_t)i}; ids[out.count++] = (uint32_t)i; }
    if (out.count < 3) return 0;
    out.faceId = faceFeature(ids, out.count);
    return out.count;
}
VXD int supportFace(const Convex& cv, V3 dir, FacePoly& out) {
    if (cv.kind == CBox) return boxFace(cv, dir, out);
    if (cv.kind == CHull) return hullFace(cv, dir, out);
    return 0;
}
// manifold.h:226-255: Sutherland-Hodgman against one clip edge.
VXD int clipEdge(const Clip2* in, int n, const Clip2& ea, const Clip2& eb, Clip2* out) {
    if (n == 0) return 0;
    const float eu = eb.u - ea.u, ev = eb.v - ea.v;
    int m = 0;
    for (int i = 0; i < n; ++i) {
        const Clip2& prev = in[(i + n - 1) % n];
        const Clip2& cur = in[i];
        float ps = eu * (prev.v - ea.v) - ev * (prev.u - ea.u);
        float cs = eu * (cur.v - ea.v) - ev * (cur.u - ea.u);
        bool pin = ps >= -1e-6f, cin = cs >= -1e-6f;
        if (pin != cin && m < 64) {
            float den = ps - cs;
            float t = fabsf(den) > 1e-12f ? ps / den : 0.0f;
            t = fclamp(t, 0.0f, 1.0f);
            out[m++] = {prev.u + (cur.u - prev.u) * t, prev.v + (cur.v - prev.v) * t,
                        edgeFeature(prev.fa, cur.fa), edgeFeature(ea.fb, eb.fb)};
        }
        if (cin && m < 64) out[m++] = cur;
    }
    return m;
}
// manifold.h:257-279
VXD int clipFaces(const FacePoly& sub, const FacePoly& clip, V3 n, Clip2* out, V3& origin, V3& u, V3& v) {
    planeBasis(n, u, v);
    origin = clip.v[0].p;
    Clip2 cur[64], tmp[64], c2[64];
    for (int i = 0; i < sub.count; ++i) {
        V3 d = sub.v[i].p - origin;
        cur[i] = {dot(d, u), dot(d, v), sub.v[i].id, clip.faceId};
    }
    for (int i = 0; i < clip.count; ++i) {
        V3 d = clip.v[i].p - origin;
        c2[i] = {dot(d, u), dot(d, v), 0, clip.v[i].id};
    }
    int count = sub.count;
    for (int i = 0; i < clip.count && count > 0; ++i) {
        count = clipEdge(cur, count, c2[i], c2[(i + 1) % clip.count], tmp);
        for (int j = 0; j < count; ++j) cur[j] = tmp[j];
    }
    for (int i = 0; i < count; ++i) out[i] = cur[i];
    return count;
}
// manifold.h:281-317: keep the deepest point, then farthest-point selection.
VXD bool featureLess(const MPoint& a, const MPoint& b) {
    return a.fa < b.fa || (a.fa == b.fa && a.fb < b.fb);
}
VXD void reducePoints(MPoint* pts, int& count) {
    if (count <= 8) return;
    MPoint sel[8];
    bool used[64] = {};
    int deep = 0;
    for (int i = 1; i < count; ++i)
        if (pts[i].gap < pts[deep].gap || (pts[i].gap == pts[deep].gap && featureLess(pts[i], pts[deep]))) deep = i;
    sel[0] = pts[deep];
    used[deep] = true;
    for (int o = 1; o < 8; ++o) {
        int best = -1;
        float bestD = -1.0f;
        for (int i = 0; i < count; ++i) {
            if (used[i]) continue;
            float nearest = 1e30f;
            for (int j = 0; j < o; ++j) nearest = fmn(nearest, lengthSq(pts[i].pointA - sel[j].pointA));
            if (nearest > bestD || (nearest == bestD && best >= 0 && featureLess(pts[i], pts[best]))) {
                best = i; bestD = nearest;
            }
        }
        if (best < 0) break;
        sel[o] = pts[best];
        used[best] = true;
    }
    for (int i = 0; i < 8; ++i) pts[i] = sel[i];
    count = 8;
}
// manifold.h:323-367
VXD int clipFaceManifold(const Convex& A, const Convex& B, V3 n, MPoint* out) {
    FacePoly fa, fb;
    if (supportFace(A, -n, fa) < 3 || supportFace(B, n, fb) < 3) return 0;
    V3 u, v;
    planeBasis(n, u, v);
    sortCcw(fa.v, fa.count, u, v);
    sortCcw(fb.v, fb.count, u, v);
    fa.plane = dot(n, fa.v[0].p);
    fb.plane = dot(n, fb.v[0].p);
    Clip2 clipped[64];
    V3 origin;
    int count = clipFaces(fa, fb, n, clipped, origin, u, v);
    if (count == 0) return 0;
    float op = dot(n, origin);
    float gap = fa.plane - fb.plane;
    MPoint cand[64];
    int pc = count < 64 ? count : 64;
    for (int i = 0; i < pc; ++i) {
        V3 tp = origin + u * clipped[i].u + v * clipped[i].v;
        cand[i].normal = n;
        cand[i].pointA = tp + n * (fa.plane - op);
        cand[i].pointB = tp + n * (fb.plane - op);
        cand[i].gap = gap;
        cand[i].fa = clipped[i].fa;
        cand[i].fb = clipped[i].fb;
    }
    reducePoints(cand, pc);
    for (int i = 0; i < pc; ++i) out[i] = cand[i];
    return pc;
}

// ---------------------------------------------------------- contact emission --
constexpr int kMaxContactsPerPair = 12;        // narrowphase.h:12
constexpr uint32_t kNoFeature = 0xffffffffu;   // narrowphase.h:402-405
constexpr uint32_t kFaceFeature = 0x40000000u;
constexpr uint32_t kImplicitFeature = 0x80000000u;
constexpr uint32_t kTriangleFeature = 0xc0000000u;

VXD uint64_t featureKey(uint32_t a, uint32_t b) {  // narrowphase.h:414-418
    uint64_t x = a == kNoFeature ? 0 : (uint64_t)a + 1;
    uint64_t y = b == kNoFeature ? 0 : (uint64_t)b + 1;
    return (x << 32) | y;
}
VXD uint64_t clipKey(uint32_t a, uint32_t b) {   // manifold.h:369-371
    return ((uint64_t(a) + 1) << 32) | (uint64_t(b) + 1);
}

VXD V3 pointVelocity(const vx_body& b, V3 p) {  // narrowphase.h:14-20
    return v3(b.velocity) + cross(v3(b.angularVelocity), p - v3(b.position));
}
VXD V3 anchorWorld(const vx_body& b, V3 local) {  // narrowphase.h:22-25
    return v3(b.position) + (b.shape == VX_SHAPE_SPHERE ? local : rotate(q4(b.orientation), local));
}
VXD V3 anchorLocal(const vx_body& b, V3 world) {  // narrowphase.h:27-30
    V3 off = world - v3(b.position);
    return b.shape == VX_SHAPE_SPHERE ? off : rotateInv(q4(b.orientation), off);
}
VXD float liveGap(const vx_body& a, const vx_body& b, const vx_contact& c) {  // narrowphase.h:32-36
    V3 pa = anchorWorld(a, v3(c.localAnchorA)), pb = anchorWorld(b, v3(c.localAnchorB));
    return c.bias0 + dot(v3(c.normal), pa - pb);
}
// narrowphase.h:38-50
VXD void tangents(V3 n, V3& t1, V3& t2) {
    V3 an{fabsf(n.x), fabsf(n.y), fabsf(n.z)};
    V3 ref = an.x < an.y ? (an.x < an.z ? V3{1, 0, 0} : V3{0, 0, 1})
                         : (an.y < an.z ? V3{0, 1, 0} : V3{0, 0, 1});
    t1 = normalize(cross(n, ref));
    t2 = cross(n, t1);
}
// narrowphase.h:52-66
VXD float combine(float a, float b, uint8_t ma, uint8_t mb) {
    uint8_t m = ma >= mb ? ma : mb;
    switch (m) {
    case VX_COMBINE_AVERAGE: return 0.5f * a + 0.5f * b;
    case VX_COMBINE_GEOMETRIC_MEAN: return sqrtf(a) * sqrtf(b);
    case VX_COMBINE_MINIMUM: return fmn(a, b);
    case VX_COMBINE_MULTIPLY: return b > 0.0f && a > 3.402823466e+38F / b ? 3.402823466e+38F : a * b;
    case VX_COMBINE_MAXIMUM: return fmx(a, b);
    }
    return 0.0f;
}
// narrowphase.h:68-75
VXD float directionalScale(const vx_body& b, V3 dir) {
    V3 l = rotateInv(q4(b.orientation), dir);
    V3 s{l.x * b.frictionScale.x, l.y * b.frictionScale.y, l.z * b.frictionScale.z};
    return length(s);
}

// narrowphase.h:425-463
VXD void emit(const vx_body& a, const vx_body& b, uint32_t ia, uint32_t ib, V3 n, V3 p, float gap,
                 float dt, vx_contact* out, int cap, int& count, uint64_t key = 0) {
    if (count >= cap) return;
    float reach = (sweeps(a) ? maxPointSpeed(a) : 0.0f) * dt + (sweeps(b) ? maxPointSpeed(b) : 0.0f) * dt +
                  a.ccdTuning.speculativeDistance + b.ccdTuning.speculativeDistance + 1e-3f;
    if (gap > reach) return;
    float vn = dot(pointVelocity(a, p) - pointVelocity(b, p), n);
    V3 la = anchorLocal(a, p), lb = anchorLocal(b, p);
    V3 t1, t2;
    tangents(n, t1, t2);
    float fr = combine(a.friction, b.friction, a.frictionCombine, b.frictionCombine);
    float s1 = sqrtf(directionalScale(a, t1) * directionalScale(b, t1));
    float s2 = sqrtf(directionalScale(a, t2) * directionalScale(b, t2));
    float rest = combine(a.restitution, b.restitution, a.restitutionCombine, b.restitutionCombine);
    float roll = combine(a.rollingFriction, b.rollingFriction, a.frictionCombine, b.frictionCombine);
    float spin = combine(a.spinningFriction, b.spinningFriction, a.frictionCombine, b.frictionCombine);
    vx_contact c{};
    c.a = ia; c.b = ib; c.featureKey = key;
    c.normal = raw(n); c.point = raw(p); c.localAnchorA = raw(la); c.localAnchorB = raw(lb);
    c.gap = gap; c.bias0 = gap; c.vn0 = vn; c.restitution = rest;
    c.friction1 = fr * s1; c.friction2 = fr * s2;
    c.rollingFriction = roll; c.spinningFriction = spin;
    out[count++] = c;
}

VXD bool convexVolume(uint8_t s) {  // narrowphase.h:646-651
    return s == VX_SHAPE_SPHERE || s == VX_SHAPE_BOX || s == VX_SHAPE_CAPSULE || s == VX_SHAPE_HULL ||
           s == VX_SHAPE_CYLINDER || s == VX_SHAPE_CONE || s == VX_SHAPE_ROUNDED_BOX || s == VX_SHAPE_ELLIPSOID;
}

// narrowphase.h:465-546
VXD void planeConvex(const vx_body& cv, const vx_body& pl, uint32_t ic, uint32_t ip, const Soup& soup,
                        float dt, vx_contact* out, int cap, int& n) {
    const V3 pn = v3(pl.planeNormal);
    const V3 pos = v3(cv.position);
    const Q4 q = q4(cv.orientation);
    switch (cv.shape) {
    case VX_SHAPE_HULL: {
        const vx_vec3* pts = soup.hullPoints + cv.hullFirst;
        uint32_t taken[4] = {};
        int nt = 0;
        for (int pick = 0; pick < 4 && pick < (int)cv.hullCount; ++pick) {
            float bg = 1e30f; V3 bp{}; uint32_t bi = 0;
            for (uint32_t i = 0; i < cv.hullCount; ++i) {
                V3 p = pos + rotate(q, v3(pts[i]));
                float g = dot(pn, p) - pl.planeOffset;
                bool used = false;
                for (int k = 0; k < nt; ++k) if (taken[k] == i) used = true;
                if (!used && g < bg) { bg = g; bp = p; bi = i; }
            }
            taken[nt++] = bi;
            emit(cv, pl, ic, ip, pn, bp, bg, dt, out, cap, n, featureKey(bi, kFaceFeature));
        }
        break;
    }
    case VX_SHAPE_SPHERE:
        emit(cv, pl, ic, ip, pn, pos - pn * cv.radius, dot(pn, pos) - pl.planeOffset - cv.radius, dt, out, cap, n,
             featureKey(kImplicitFeature, kFaceFeature));
        break;
    case VX_SHAPE_CAPSULE: {
        V3 ax = rotate(q, V3{0, cv.capsuleHalfHeight, 0});
        V3 ends[2] = {pos + ax, pos - ax};
        for (int e = 0; e < 2; ++e)
            emit(cv, pl, ic, ip, pn, ends[e] - pn * cv.radius, dot(pn, ends[e]) - pl.planeOffset - cv.radius, dt,
                 out, cap, n, featureKey(kImplicitFeature | (uint32_t)e, kFaceFeature));
        break;
    }
    case VX_SHAPE_BOX: {
        V3 pts[8]; float gaps[8]; uint32_t ft[8];
        for (int i = 0; i < 8; ++i) {
            V3 l{(i & 1) ? cv.halfExtents.x : -cv.halfExtents.x, (i & 2) ? cv.halfExtents.y : -cv.halfExtents.y,
                 (i & 4) ? cv.halfExtents.z : -cv.halfExtents.z};
            pts[i] = pos + rotate(q, l);
            gaps[i] = dot(pn, pts[i]) - pl.planeOffset;
            ft[i] = (uint32_t)i;
        }
        for (int pick = 0; pick < 4; ++pick) {
            int best = pick;
            for (int i = pick + 1; i < 8; ++i) if (gaps[i] < gaps[best]) best = i;
            float tg = gaps[pick]; gaps[pick] = gaps[best]; gaps[best] = tg;
            V3 tp = pts[pick]; pts[pick] = pts[best]; pts[best] = tp;
            uint32_t tf = ft[pick]; ft[pick] = ft[best]; ft[best] = tf;
            emit(cv, pl, ic, ip, pn, pts[pick], gaps[pick], dt, out, cap, n, featureKey(ft[pick], kFaceFeature));
        }
        break;
    }
    case VX_SHAPE_CYLINDER:
    case VX_SHAPE_CONE: {
        Convex c = makeConvex(cv, soup);
        V3 deep = c.support(-pn);
        emit(cv, pl, ic, ip, pn, deep, dot(pn, deep) - pl.planeOffset, dt, out, cap, n,
             featureKey(kImplicitFeature, kFaceFeature));
        break;
    }
    default: break;
    }
}

// narrowphase.h:555-607
VXD void convexConvex(const vx_body& a, const vx_body& b, uint32_t ia, uint32_t ib, const Soup& soup,
                         float dt, vx_contact* out, int cap, int& n) {
    Convex ca = makeConvex(a, soup), cb = makeConvex(b, soup);
    GjkResult r = gjkDistance(ca, cb);
    const bool flatA = a.shape == VX_SHAPE_BOX || a.shape == VX_SHAPE_HULL;
    const bool flatB = b.shape == VX_SHAPE_BOX || b.shape == VX_SHAPE_HULL;
    if (r.distance < 0.05f && flatA && flatB) {
        V3 mn = r.normal;
        if (a.shape == VX_SHAPE_BOX && b.shape == VX_SHAPE_BOX) {
            float best = -1e30f;
            for (int side = 0; side < 2; ++side) {
                const vx_body& box = side ? b : a;
                for (int ax = 0; ax < 3; ++ax) {
                    V3 cand = rotate(q4(box.orientation), V3{ax == 0 ? 1.0f : 0.0f, ax == 1 ? 1.0f : 0.0f,
                                                             ax == 2 ? 1.0f : 0.0f});
                    if (dot(cand, v3(a.position) - v3(b.position)) < 0.0f) cand = -cand;
                    float sep = dot(cand, ca.support(-cand) - cb.support(cand));
                    if (sep > best) { best = sep; mn = cand; }
                }
            }
        }
        MPoint mp[8];
        int pc = clipFaceManifold(ca, cb, mn, mp);
        if (pc > 0) {
            bool emitted = false;
            for (int i = 0; i < pc && n < cap; ++i) {
                emit(a, b, ia, ib, mp[i].normal, (mp[i].pointA + mp[i].pointB) * 0.5f, mp[i].gap, dt, out, cap, n,
                     clipKey(mp[i].fa, mp[i].fb));
                emitted = true;
            }
            if (emitted) return;
        }
    }
    emit(a, b, ia, ib, r.normal, (r.pointA + r.pointB) * 0.5f, r.distance, dt, out, cap, n);
}

// narrowphase.h:369-398
VXD void bodyAabb(const vx_body& b, float dt, V3& lo, V3& hi) {
    float ext;
    const V3 he = v3(b.halfExtents);
    switch (b.shape) {
    case VX_SHAPE_BOX: ext = length(he); break;
    case VX_SHAPE_ROUNDED_BOX: ext = length(he) + b.radius; break;
    case VX_SHAPE_ELLIPSOID: ext = fmx(he.x, fmx(he.y, he.z)); break;
    case VX_SHAPE_CAPSULE: ext = b.capsuleHalfHeight + b.radius; break;
    case VX_SHAPE_SPHERE: case VX_SHAPE_HULL: case VX_SHAPE_COMPOUND: ext = b.radius; break;
    case VX_SHAPE_CYLINDER: ext = sqrtf(b.radius * b.radius + b.capsuleHalfHeight * b.capsuleHalfHeight); break;
    case VX_SHAPE_CONE: {
        float top = 1.5f * b.capsuleHalfHeight;
        ext = fmx(top, sqrtf(b.radius * b.radius + 0.25f * b.capsuleHalfHeight * b.capsuleHalfHeight));
        break;
    }
    default: ext = 0.0f; break;
    }
    const float sweep = sweeps(b) ? maxPointSpeed(b) * dt : 0.0f;
    const float reach = ext + sweep + b.ccdTuning.collisionMargin + 1e-2f;
    lo = v3(b.position) - V3{reach, reach, reach};
    hi = v3(b.position) + V3{reach, reach, reach};
}
VXD bool aabbOverlap(V3 l1, V3 h1, V3 l2, V3 h2) {  // narrowphase.h:400-405
    return h1.x >= l2.x && l1.x <= h2.x && h1.y >= l2.y && l1.y <= h2.y && h1.z >= l2.z && l1.z <= h2.z;
}

// narrowphase.h:609-643
VXD void meshConvex(const vx_body& cv, const vx_body& mb, uint32_t ic, uint32_t im, const Soup& soup,
                       float dt, vx_contact* out, int cap, int& n) {
    const vx_mesh& m = soup.meshes[mb.meshIndex];
    V3 lo, hi;
    bodyAabb(cv, dt, lo, hi);
    Convex cc = makeConvex(cv, soup);
    uint32_t stack[48];
    int sp = 0;
    stack[sp++] = m.firstNode;
    while (sp > 0) {
        const vx_bvh_node& nd = soup.nodes[stack[--sp]];
        if (!aabbOverlap(lo, hi, v3(nd.aabbMin), v3(nd.aabbMax))) continue;
        if (nd.triCount == 0) {
            if (sp + 2 <= 48) { stack[sp++] = nd.leftFirst; stack[sp++] = nd.leftFirst + 1; }
            continue;
        }
        for (uint32_t k = 0; k < nd.triCount; ++k) {
            uint32_t tri = soup.triRefs[nd.leftFirst + k];
            uint32_t base = m.firstIndex + tri * 3;
            GjkResult r = gjkDistance(cc, makeTriangle(v3(soup.vertices[m.firstVertex + soup.indices[base]]),
                                                       v3(soup.vertices[m.firstVertex + soup.indices[base + 1]]),
                                                       v3(soup.vertices[m.firstVertex + soup.indices[base + 2]])));
            emit(cv, mb, ic, im, r.normal, (r.pointA + r.pointB) * 0.5f, r.distance, dt, out, cap, n,
                 featureKey(kImplicitFeature, kTriangleFeature | tri));
        }
    }
}

// narrowphase.h:656-673
VXD int collideSimple(const vx_body& a, const vx_body& b, uint32_t ia, uint32_t ib, const Soup& soup,
                         float dt, vx_contact* out, int cap) {
    int n = 0;
    if (convexVolume(a.shape) && convexVolume(b.shape)) convexConvex(a, b, ia, ib, soup, dt, out, cap, n);
    else if (convexVolume(a.shape) && b.shape == VX_SHAPE_PLANE) planeConvex(a, b, ia, ib, soup, dt, out, cap, n);
    else if (a.shape == VX_SHAPE_PLANE && convexVolume(b.shape)) planeConvex(b, a, ib, ia, soup, dt, out, cap, n);
    else if (convexVolume(a.shape) && b.shape == VX_SHAPE_MESH) meshConvex(a, b, ia, ib, soup, dt, out, cap, n);
    else if (a.shape == VX_SHAPE_MESH && convexVolume(b.shape)) meshConvex(b, a, ib, ia, soup, dt, out, cap, n);
    return n;
}

// backend.h:42-57
VXD vx_body childBody(const vx_body& parent, const vx_compound_child& ch) {
    vx_body r = parent;
    r.shape = ch.shape;
    r.position = raw(v3(parent.position) + rotate(q4(parent.orientation), v3(ch.localPosition)));
    r.orientation = raw(qmul(q4(parent.orientation), q4(ch.localOrientation)));
    r.radius = ch.radius;
    r.halfExtents = ch.halfExtents;
    r.capsuleHalfHeight = ch.capsuleHalfHeight;
    r.hullFirst = ch.hullFirst; r.hullCount = ch.hullCount;
    r.hullFaceFirst = ch.hullFaceFirst; r.hullFaceCount = ch.hullFaceCount;
    r.compoundFirst = r.compoundCount = 0;
    return r;
}

// narrowphase.h:678-714
VXD int collidePair(const vx_body& a, const vx_body& b, uint32_t ia, uint32_t ib, const Soup& soup, float dt,
                       vx_contact* out, int cap) {
    if (a.shape != VX_SHAPE_COMPOUND && b.shape != VX_SHAPE_COMPOUND)
        return collideSimple(a, b, ia, ib, soup, dt, out, cap);
    int na = a.shape == VX_SHAPE_COMPOUND ? (int)a.compoundCount : 1;
    int nb = b.shape == VX_SHAPE_COMPOUND ? (int)b.compoundCount : 1;
    int total = 0;
    for (int x = 0; x < na && total < cap; ++x) {
        vx_body ga = a.shape == VX_SHAPE_COMPOUND ? childBody(a, soup.children[a.compoundFirst + x]) : a;
        for (int y = 0; y < nb && total < cap; ++y) {
            vx_body gb = b.shape == VX_SHAPE_COMPOUND ? childBody(b, soup.children[b.compoundFirst + y]) : b;
            int added = collideSimple(ga, gb, ia, ib, soup, dt, out + total, cap - total);
            for (int k = 0; k < added; ++k) {
                vx_contact& c = out[total + k];
                const vx_body& oa = c.a == ia ? a : b;
                const vx_body& ob = c.b == ib ? b : a;
                V3 p = v3(c.point);
                c.localAnchorA = raw(rotateInv(q4(oa.orientation), p - v3(oa.position)));
                c.localAnchorB = raw(rotateInv(q4(ob.orientation), p - v3(ob.position)));
                c.vn0 = dot(pointVelocity(oa, p) - pointVelocity(ob, p), v3(c.normal));
                uint64_t tag = (uint64_t)(x + 1) * 0x9e3779b185ebca87ull ^ (uint64_t)(y + 1) * 0xc2b2ae3d27d4eb4full;
                c.featureKey ^= tag;
            }
            total += added;
        }
    }
    return total;
}

// narrowphase.h:719-764: minimum gap to a mesh inside a search box.
VXD void meshGapProbe1(const vx_body& cv, const vx_body& mb, const Soup& soup, float search, float& bestGap,
                          V3& bestN) {
    const vx_mesh& m = soup.meshes[mb.meshIndex];
    V3 r{search, search, search};
    V3 lo = v3(cv.position) - r, hi = v3(cv.position) + r;
    Convex cc = makeConvex(cv, soup);
    uint32_t stack[48];
    int sp = 0;
    stack[sp++] = m.firstNode;
    while (sp > 0) {
        const vx_bvh_node& nd = soup.nodes[stack[--sp]];
        if (!aabbOverlap(lo, hi, v3(nd.aabbMin), v3(nd.aabbMax))) continue;
        if (nd.triCount == 0) {
            if (sp + 2 <= 48) { stack[sp++] = nd.leftFirst; stack[sp++] = nd.leftFirst + 1; }
            continue;
        }
        for (uint32_t k = 0; k < nd.triCount; ++k) {
            uint32_t tri = soup.triRefs[nd.leftFirst + k];
            uint32_t base = m.firstIndex + tri * 3;
            GjkResult g = gjkDistance(cc, makeTriangle(v3(soup.vertices[m.firstVertex + soup.indices[base]]),
                                                       v3(soup.vertices[m.firstVertex + soup.indices[base + 1]]),
                                                       v3(soup.vertices[m.firstVertex + soup.indices[base + 2]])));
            if (g.distance < bestGap) { bestGap = g.distance; bestN = g.normal; }
        }
    }
}
VXD void meshGapProbe(const vx_body& cv, const vx_body& mb, const Soup& soup, float search, float& bestGap,
                         V3& bestN) {
    if (cv.shape != VX_SHAPE_COMPOUND) { meshGapProbe1(cv, mb, soup, search, bestGap, bestN); return; }
    for (uint32_t i = 0; i < cv.compoundCount; ++i)
        meshGapProbe1(childBody(cv, soup.children[cv.compoundFirst + i]), mb, soup, search, bestGap, bestN);
}

}  // namespace vxd
