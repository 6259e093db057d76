// Host-side scene builder: turns a vxscene::Scene description into Velox body
// records and flat geometry (vx_body / vx_soup) for the device step.
//
// Closed-form inertia of spheres, boxes and capsules follows the reference's
// creation functions (world.cpp:741-823); hulls get facets by exhaustive
// supporting-plane enumeration (unique for points in general position), mass
// properties by signed-tetrahedron integration in double precision and
// principal axes by cyclic Jacobi rotations; heightfields are triangulated and
// given a median-split triangle BVH in the layout of world.cpp:1320-1477
// (leaves of at most 4 triangles, right child = left + 1).
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>
#include "../../include/vxhip.h"
#include "scene_gen.hpp"

struct vxhip_scene {
    std::vector<vx_body> bodies;
    std::vector<vx_vec3> vertices, hullPoints;
    std::vector<uint32_t> indices, triRefs, hullFaces;
    std::vector<vx_mesh> meshes;
    std::vector<vx_bvh_node> nodes;
    float gravity[3] = {0, -9.81f, 0};
    float dt = 1.0f / 60.0f;
    int substeps = 4;
    std::vector<uint32_t> worldOf;   // batched scenes: owning world of each body
    uint32_t worlds = 1;
    std::vector<vx_joint> joints;
};

namespace {

vx_body defaultBody() {  // body.h:128-171 defaults + World::addBody CCD defaults (world.cpp:646-652)
    vx_body b;
    std::memset(&b, 0, sizeof(b));
    b.orientation = {0, 0, 0, 1};
    b.invMass = 1.0f;
    b.inertiaOrientation = {0, 0, 0, 1};
    b.restitution = 0.3f;
    b.friction = 0.5f;
    b.frictionScale = {1, 1, 1};
    b.frictionCombine = VX_COMBINE_GEOMETRIC_MEAN;
    b.restitutionCombine = VX_COMBINE_MINIMUM;
    b.gravityScale = 1.0f;
    b.motionType = VX_MOTION_DYNAMIC;
    b.ccdTuning.quality = VX_QUALITY_MEDIUM;
    b.ccdTuning.collisionMargin = 0.0f;
    b.ccdTuning.speculativeDistance = 0.01f;
    b.ccdTuning.enableContinuous = 1;
    b.ccdTuning.minVelocityForCCD = 1.0f;
    vx_ccd_settings& s = b.ccdTuning.ccdSettings;  // CcdQuality::Balanced preset (ccd.h:53-63)
    s.quality = 1; s.maxAdvancementIterations = 16; s.maxRootFindIterations = 12;
    s.distanceTolerance = 1e-3f; s.timeTolerance = 1e-5f; s.enableRotationalCcd = 1; s.enableCompoundCcd = 1;
    s.useBrentRootFinder = 1; s.meshBvhMaxDepth = 8; s.speculativeMargin = 0.01f;
    b.categoryBits = 1u;
    b.maskBits = 0xffffffffu;
    b.enableSleep = 1;
    b.shape = VX_SHAPE_SPHERE;
    b.radius = 0.5f;
    b.halfExtents = {0.5f, 0.5f, 0.5f};
    b.capsuleHalfHeight = 0.5f;
    b.planeNormal = {0, 1, 0};
    return b;
}

void setMass(vx_body& b, float mass) {
    b.motionType = mass > 0.0f ? VX_MOTION_DYNAMIC : VX_MOTION_STATIC;
    b.invMass = mass > 0.0f ? 1.0f / mass : 0.0f;
    b.invInertia = {0, 0, 0};
}

struct D3 { double x, y, z; };
D3 sub(D3 a, D3 b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
D3 crs(D3 a, D3 b) { return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x}; }
double dt3(D3 a, D3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }

// Facets of the convex hull of points in general position: every triple whose
// plane leaves all other points on one side, wound outwards.
std::vector<uint32_t> hullFacets(const std::vector<D3>& p) {
    const uint32_t n = (uint32_t)p.size();
    D3 c{0, 0, 0};
    for (const D3& q : p) { c.x += q.x / n; c.y += q.y / n; c.z += q.z / n; }
    double scale = 0.0;
    for (const D3& q : p) scale = std::max(scale, std::sqrt(dt3(sub(q, c), sub(q, c))));
    const double eps = 1e-9 * scale * scale * scale;
    std::vector<uint32_t> tris;
    for (uint32_t i = 0; i < n; ++i)
        for (uint32_t j = i + 1; j < n; ++j)
            for (uint32_t k = j + 1; k < n; ++k) {
                D3 nrm = crs(sub(p[j], p[i]), sub(p[k], p[i]));
                int pos = 0, neg = 0;
                for (uint32_t m = 0; m < n && !(pos && neg); ++m) {
                    if (m == i || m == j || m == k) continue;
                    double d = dt3(nrm, sub(p[m], p[i]));
                    if (d > eps) ++pos; else if (d < -eps) ++neg;
                }
                if (pos && neg) continue;
                if (pos) tris.insert(tris.end(), {i, k, j});  // others above: flip outwards
                else tris.insert(tris.end(), {i, j, k});
            }
    return tris;
}

// Cyclic Jacobi diagonalisation of a symmetric 3x3; columns of V are the axes.
void jacobi(double a[3][3], double eig[3], double V[3][3]) {
    for (int i = 0; i < 3; ++i) for (int j = 0; j < 3; ++j) V[i][j] = i == j ? 1.0 : 0.0;
    for (int sweep = 0; sweep < 64; ++sweep) {
        double off = a[0][1] * a[0][1] + a[0][2] * a[0][2] + a[1][2] * a[1][2];
        if (off < 1e-30) break;
        for (int p = 0; p < 2; ++p)
            for (int q = p + 1; q < 3; ++q) {
                if (std::fabs(a[p][q]) < 1e-300) continue;
                double theta = (a[q][q] - a[p][p]) / (2.0 * a[p][q]);
                double t = (theta >= 0 ? 1.0 : -1.0) / (std::fabs(theta) + std::sqrt(theta * theta + 1.0));
                double cs = 1.0 / std::sqrt(t * t + 1.0), sn = t * cs;
                for (int k = 0; k < 3; ++k) {
                    double akp = a[k][p], akq = a[k][q];
                    a[k][p] = cs * akp - sn * akq;
                    a[k][q] = sn * akp + cs * akq;
                }
                for (int k = 0; k < 3; ++k) {
                    double apk = a[p][k], aqk = a[q][k];
                    a[p][k] = cs * apk - sn * aqk;
                    a[q][k] = sn * apk + cs * aqk;
                }
                for (int k = 0; k < 3; ++k) {
                    double vkp = V[k][p], vkq = V[k][q];
                    V[k][p] = cs * vkp - sn * vkq;
                    V[k][q] = sn * vkp + cs * vkq;
                }
            }
    }
    for (int i = 0; i < 3; ++i) eig[i] = a[i][i];
}

vx_quat quatFromColumns(double V[3][3]) {
    // Right-handed frame: flip the third axis if needed.
    D3 c0{V[0][0], V[1][0], V[2][0]}, c1{V[0][1], V[1][1], V[2][1]}, c2{V[0][2], V[1][2], V[2][2]};
    if (dt3(crs(c0, c1), c2) < 0) { V[0][2] = -V[0][2]; V[1][2] = -V[1][2]; V[2][2] = -V[2][2]; }
    double m00 = V[0][0], m11 = V[1][1], m22 = V[2][2], tr = m00 + m11 + m22;
    double x, y, z, w;
    if (tr > 0) {
        double s = std::sqrt(tr + 1.0) * 2;
        w = 0.25 * s; x = (V[2][1] - V[1][2]) / s; y = (V[0][2] - V[2][0]) / s; z = (V[1][0] - V[0][1]) / s;
    } else if (m00 > m11 && m00 > m22) {
        double s = std::sqrt(1.0 + m00 - m11 - m22) * 2;
        w = (V[2][1] - V[1][2]) / s; x = 0.25 * s; y = (V[0][1] + V[1][0]) / s; z = (V[0][2] + V[2][0]) / s;
    } else if (m11 > m22) {
        double s = std::sqrt(1.0 + m11 - m00 - m22) * 2;
        w = (V[0][2] - V[2][0]) / s; x = (V[0][1] + V[1][0]) / s; y = 0.25 * s; z = (V[1][2] + V[2][1]) / s;
    } else {
        double s = std::sqrt(1.0 + m22 - m00 - m11) * 2;
        w = (V[1][0] - V[0][1]) / s; x = (V[0][2] + V[2][0]) / s; y = (V[1][2] + V[2][1]) / s; z = 0.25 * s;
    }
    double len = std::sqrt(x * x + y * y + z * z + w * w);
    return {(float)(x / len), (float)(y / len), (float)(z / len), (float)(w / len)};
}

void addHull(vxhip_scene& sc, vx_body& b, const float* pts, uint32_t count, float mass) {
    std::vector<D3> p(count);
    for (uint32_t i = 0; i < count; ++i) p[i] = {pts[3 * i], pts[3 * i + 1], pts[3 * i + 2]};
    std::vector<uint32_t> tris = hullFacets(p);
    double vol = 0, first[3] = {0, 0, 0}, second[3][3] = {};
    for (size_t t = 0; t < tris.size(); t += 3) {
        const D3 &a = p[tris[t]], &bb = p[tris[t + 1]], &c = p[tris[t + 2]];
        double v = dt3(a, crs(bb, c)) / 6.0;
        vol += v;
        const double q[3][3] = {{a.x, a.y, a.z}, {bb.x, bb.y, bb.z}, {c.x, c.y, c.z}};
        for (int ax = 0; ax < 3; ++ax) first[ax] += v * (q[0][ax] + q[1][ax] + q[2][ax]) / 4.0;
        for (int x = 0; x < 3; ++x)
            for (int y = 0; y < 3; ++y) {
                double s = 0;
                for (int i = 0; i < 3; ++i) s += 2.0 * q[i][x] * q[i][y];
                for (int i = 0; i < 3; ++i) for (int j = 0; j < 3; ++j) if (i != j) s += q[i][x] * q[j][y];
                second[x][y] += v * s / 20.0;
            }
    }
    double ctr[3] = {first[0] / vol, first[1] / vol, first[2] / vol};
    double I[3][3];
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) {
            double origin = i == j ? second[(i + 1) % 3][(i + 1) % 3] + second[(i + 2) % 3][(i + 2) % 3] : -second[i][j];
            double shift = vol * ((i == j ? ctr[0] * ctr[0] + ctr[1] * ctr[1] + ctr[2] * ctr[2] : 0.0) - ctr[i] * ctr[j]);
            I[i][j] = origin - shift;
        }
    double eig[3], V[3][3];
    jacobi(I, eig, V);
    const vx_vec3 center{(float)ctr[0], (float)ctr[1], (float)ctr[2]};
    b.shape = VX_SHAPE_HULL;
    setMass(b, mass);
    if (mass > 0.0f) {
        b.position.x += center.x; b.position.y += center.y; b.position.z += center.z;
    }
    b.hullFirst = (uint32_t)sc.hullPoints.size();
    b.hullCount = count;
    b.hullFaceFirst = (uint32_t)sc.hullFaces.size();
    b.hullFaceCount = (uint32_t)(tris.size() / 3);
    float r2 = 0.0f;
    for (uint32_t i = 0; i < count; ++i) {
        vx_vec3 q{pts[3 * i], pts[3 * i + 1], pts[3 * i + 2]};
        if (mass > 0.0f) { q.x -= center.x; q.y -= center.y; q.z -= center.z; }
        sc.hullPoints.push_back(q);
        r2 = std::max(r2, q.x * q.x + q.y * q.y + q.z * q.z);
    }
    sc.hullFaces.insert(sc.hullFaces.end(), tris.begin(), tris.end());
    b.radius = std::sqrt(r2);
    if (mass > 0.0f) {
        float density = mass / (float)vol;
        b.invInertia = {1.0f / ((float)eig[0] * density), 1.0f / ((float)eig[1] * density),
                        1.0f / ((float)eig[2] * density)};
        b.inertiaOrientation = quatFromColumns(V);
    }
}

struct BvhBuild {
    std::vector<vx_bvh_node>& nodes;
    std::vector<uint32_t>& refs;
    const std::vector<vx_vec3>& cen;
    const std::vector<vx_vec3>& tmin;
    const std::vector<vx_vec3>& tmax;
    void bounds(uint32_t node, uint32_t first, uint32_t count) {
        vx_bvh_node& nd = nodes[node];
        nd.aabbMin = tmin[refs[first]];
        nd.aabbMax = tmax[refs[first]];
        for (uint32_t i = 1; i < count; ++i) {
            const vx_vec3 &lo = tmin[refs[first + i]], &hi = tmax[refs[first + i]];
            nd.aabbMin = {std::min(nd.aabbMin.x, lo.x), std::min(nd.aabbMin.y, lo.y), std::min(nd.aabbMin.z, lo.z)};
            nd.aabbMax = {std::max(nd.aabbMax.x, hi.x), std::max(nd.aabbMax.y, hi.y), std::max(nd.aabbMax.z, hi.z)};
        }
    }
    void build(uint32_t node, uint32_t first, uint32_t count) {
        bounds(node, first, count);
        if (count <= 4) { nodes[node].leftFirst = first; nodes[node].triCount = count; return; }
        vx_vec3 lo = cen[refs[first]], hi = lo;
        for (uint32_t i = 1; i < count; ++i) {
            const vx_vec3& c = cen[refs[first + i]];
            lo = {std::min(lo.x, c.x), std::min(lo.y, c.y), std::min(lo.z, c.z)};
            hi = {std::max(hi.x, c.x), std::max(hi.y, c.y), std::max(hi.z, c.z)};
        }
        float ex = hi.x - lo.x, ey = hi.y - lo.y, ez = hi.z - lo.z;
        int axis = ex > ey ? (ex > ez ? 0 : 2) : (ey > ez ? 1 : 2);
        auto key = [&](uint32_t t) { return axis == 0 ? cen[t].x : axis == 1 ? cen[t].y : cen[t].z; };
        std::nth_element(refs.begin() + first, refs.begin() + first + count / 2, refs.begin() + first + count,
                         [&](uint32_t a, uint32_t b) { return key(a) < key(b); });
        uint32_t left = (uint32_t)nodes.size();
        nodes[node].leftFirst = left;
        nodes[node].triCount = 0;
        nodes.emplace_back();
        nodes.emplace_back();
        build(left, first, count / 2);
        build(left + 1, first + count / 2, count - count / 2);
    }
};

void addHeightfield(vxhip_scene& sc, vx_body& b, const vxscene::HeightfieldSpec& hf) {
    std::vector<vx_vec3> v;
    std::vector<uint32_t> ix;
    for (uint32_t z = 0; z < hf.depth; ++z)
        for (uint32_t x = 0; x < hf.width; ++x)
            v.push_back({hf.origin[0] + x * hf.cell, hf.origin[1] + hf.heights[(size_t)z * hf.width + x],
                         hf.origin[2] + z * hf.cell});
    for (uint32_t z = 0; z + 1 < hf.depth; ++z)
        for (uint32_t x = 0; x + 1 < hf.width; ++x) {
            uint32_t a = z * hf.width + x, bb = a + 1, c = a + hf.width, d = c + 1;
            ix.insert(ix.end(), {a, c, bb, bb, c, d});
        }
    vx_mesh m;
    std::memset(&m, 0, sizeof(m));
    m.firstVertex = (uint32_t)sc.vertices.size();
    m.vertexCount = (uint32_t)v.size();
    m.firstIndex = (uint32_t)sc.indices.size();
    m.indexCount = (uint32_t)ix.size();
    m.aabbMin = m.aabbMax = v[0];
    for (const vx_vec3& q : v) {
        m.aabbMin = {std::min(m.aabbMin.x, q.x), std::min(m.aabbMin.y, q.y), std::min(m.aabbMin.z, q.z)};
        m.aabbMax = {std::max(m.aabbMax.x, q.x), std::max(m.aabbMax.y, q.y), std::max(m.aabbMax.z, q.z)};
    }
    sc.vertices.insert(sc.vertices.end(), v.begin(), v.end());
    sc.indices.insert(sc.indices.end(), ix.begin(), ix.end());
    const uint32_t tris = m.indexCount / 3;
    std::vector<vx_vec3> cen(tris), tmin(tris), tmax(tris);
    for (uint32_t t = 0; t < tris; ++t) {
        const vx_vec3 &a = v[ix[3 * t]], &bb = v[ix[3 * t + 1]], &c = v[ix[3 * t + 2]];
        cen[t] = {(a.x + bb.x + c.x) * (1.0f / 3.0f), (a.y + bb.y + c.y) * (1.0f / 3.0f), (a.z + bb.z + c.z) * (1.0f / 3.0f)};
        tmin[t] = {std::min(std::min(a.x, bb.x), c.x), std::min(std::min(a.y, bb.y), c.y), std::min(std::min(a.z, bb.z), c.z)};
        tmax[t] = {std::max(std::max(a.x, bb.x), c.x), std::max(std::max(a.y, bb.y), c.y), std::max(std::max(a.z, bb.z), c.z)};
    }
    m.firstNode = (uint32_t)sc.nodes.size();
    m.firstTriRef = (uint32_t)sc.triRefs.size();
    for (uint32_t t = 0; t < tris; ++t) sc.triRefs.push_back(t);
    sc.nodes.emplace_back();
    BvhBuild bvh{sc.nodes, sc.triRefs, cen, tmin, tmax};
    bvh.build(m.firstNode, m.firstTriRef, tris);
    m.nodeCount = (uint32_t)sc.nodes.size() - m.firstNode;
    sc.meshes.push_back(m);
    b.shape = VX_SHAPE_MESH;
    setMass(b, 0.0f);
    b.meshIndex = (uint32_t)sc.meshes.size() - 1;
}

void addBody(vxhip_scene& sc, const vxscene::Scene& s, const vxscene::BodySpec& sp) {
    vx_body b = defaultBody();
    b.position = {sp.pos[0], sp.pos[1], sp.pos[2]};
    const float m = sp.mass;
    switch (sp.kind) {
    case vxscene::Sphere:
        b.shape = VX_SHAPE_SPHERE; b.radius = sp.radius; setMass(b, m);
        if (m > 0.0f) { float i = 0.4f * m * sp.radius * sp.radius; b.invInertia = {1.0f / i, 1.0f / i, 1.0f / i}; }
        break;
    case vxscene::Box: {
        b.shape = VX_SHAPE_BOX; b.halfExtents = {sp.half[0], sp.half[1], sp.half[2]}; setMass(b, m);
        if (m > 0.0f) {
            float ex = sp.half[0] * 2.0f, ey = sp.half[1] * 2.0f, ez = sp.half[2] * 2.0f, k = m / 12.0f;
            b.invInertia = {1.0f / (k * (ey * ey + ez * ez)), 1.0f / (k * (ex * ex + ez * ez)), 1.0f / (k * (ex * ex + ey * ey))};
        }
        break;
    }
    case vxscene::Capsule: {
        b.shape = VX_SHAPE_CAPSULE; b.radius = sp.radius; b.capsuleHalfHeight = sp.halfHeight; setMass(b, m);
        if (m > 0.0f) {
            float r = sp.radius, hh = sp.halfHeight;
            float cw = 2.0f * hh, sw = (4.0f / 3.0f) * r;
            float cm = m * cw / (cw + sw), sm = m - cm;
            float iy = 0.5f * cm * r * r + 0.4f * sm * r * r;
            float ix = cm * (3.0f * r * r + 4.0f * hh * hh) / 12.0f + sm * (0.4f * r * r + 0.75f * hh * r + hh * hh);
            b.invInertia = {1.0f / ix, 1.0f / iy, 1.0f / ix};
        }
        break;
    }
    case vxscene::Hull: addHull(sc, b, &s.hullPoints[3 * sp.hullFirst], sp.hullCount, m); break;
    case vxscene::Plane: {
        b.shape = VX_SHAPE_PLANE; setMass(b, 0.0f);
        float l = std::sqrt(sp.normal[0] * sp.normal[0] + sp.normal[1] * sp.normal[1] + sp.normal[2] * sp.normal[2]);
        float inv = l > 1e-8f ? 1.0f / l : 0.0f;
        b.planeNormal = {sp.normal[0] * inv, sp.normal[1] * inv, sp.normal[2] * inv};
        b.planeOffset = sp.offset;
        b.position = {0, 0, 0};
        break;
    }
    case vxscene::Heightfield: b.position = {0, 0, 0}; addHeightfield(sc, b, s.heightfields[sp.hfIndex]); break;
    }
    if (sp.kind != vxscene::Plane && sp.kind != vxscene::Heightfield) {
        b.orientation = {sp.quat[0], sp.quat[1], sp.quat[2], sp.quat[3]};
        b.velocity = {sp.vel[0], sp.vel[1], sp.vel[2]};
        b.angularVelocity = {sp.angVel[0], sp.angVel[1], sp.angVel[2]};
    }
    b.friction = sp.friction;
    b.restitution = sp.restitution;
    b.enableSleep = sp.enableSleep;
    b.ccdTuning.quality = sp.quality;
    sc.bodies.push_back(b);
    sc.worldOf.push_back(sp.world);
}

}  // namespace

vxscene::Scene vxSceneByName(const std::string& name, uint32_t param, bool& ok);

// A ball joint with the reference's Joint defaults (joint.h:92-190): axes
// (0,1,0), references (1,0,0), unbreakable (FLT_MAX), no spring, motor or
// limits, all limit axes masked in, linked bodies not colliding.
vx_joint ballJoint(const vxscene::JointSpec& js) {
    vx_joint j;
    std::memset(&j, 0, sizeof(j));
    j.type = VX_JOINT_BALL;
    j.a = js.a;
    j.b = js.b;
    j.localAnchorA = {js.anchorA[0], js.anchorA[1], js.anchorA[2]};
    j.localAnchorB = {js.anchorB[0], js.anchorB[1], js.anchorB[2]};
    j.localAxisA = j.localAxisB = {0, 1, 0};
    j.localRefA = j.localRefB = {1, 0, 0};
    j.springDampingRatio = 1.0f;
    j.breakForce = j.breakTorque = 3.402823466e+38F;
    j.swingLimit = 3.14159265f;
    j.lowerTwistLimit = -3.14159265f;
    j.upperTwistLimit = 3.14159265f;
    j.linearLimitMask = j.angularLimitMask = 7;
    j.weldDampingRatio = 1.0f;
    j.suspensionFrequencyHz = 2.0f;
    j.suspensionDampingRatio = 0.7f;
    j.wheelRadius = 0.5f;
    j.maxSteeringAngle = 0.6f;
    j.pulleyRatio = 1.0f;
    j.gearRatio = 1.0f;
    return j;
}

vxhip_scene* vxBuildScene(const vxscene::Scene& s) {
    vxhip_scene* sc = new vxhip_scene();
    for (const vxscene::BodySpec& b : s.bodies) addBody(*sc, s, b);
    for (const vxscene::JointSpec& j : s.joints) sc->joints.push_back(ballJoint(j));
    std::memcpy(sc->gravity, s.gravity, sizeof(sc->gravity));
    sc->dt = s.dt;
    sc->substeps = s.substeps;
    sc->worlds = s.worlds;
    return sc;
}

vxscene::Scene vxSceneByName(const std::string& name, uint32_t param, bool& ok) {
    ok = true;
    if (name == "lattice") return vxscene::lattice(param ? (int)param : 10, param ? (int)param : 10, param ? (int)param : 10);
    if (name == "debris") return vxscene::debris(param ? param : 50000);
    if (name == "pile") {  // density of the 240-body pile kept as the count grows
        const uint32_t count = param ? param : 240;
        const float f = std::cbrt((float)count / 240.0f);
        return vxscene::pile(count, 6.0f * f, 8.0f * f);
    }
    if (name == "terrain") return vxscene::terrain(24, 0.5f, 1.0f, param ? param : 96);
    if (name == "freeflight") return vxscene::freeFlight(param ? param : 128);
    if (name == "batch") return vxscene::batchWorlds(param ? param : 4096);
    if (name == "ccdstress") return vxscene::ccdStress();
    if (name == "ccdsmall") return vxscene::ccdStressSmall();
    if (name == "settle") return vxscene::settle();
    if (name == "hightoi") return vxscene::highToi();
    if (name == "slabstack") return vxscene::slabStack();
    if (name == "diff") return vxscene::diffWorlds(param ? param : 1024);
    // The exact descriptions behind the golden trajectory fixtures.
    if (name == "fixture-pile") return vxscene::pile(160, 5.0f, 6.0f);
    if (name == "fixture-terrain") return vxscene::terrain(16, 0.5f, 0.8f, 40);
    ok = false;
    return {};
}

extern "C" {

int vxhip_scene_build(const char* name, uint32_t param, vxhip_scene** out) {
    if (!name || !out) return VXHIP_ERR_INVALID;
    *out = nullptr;
    bool ok = false;
    vxscene::Scene s = vxSceneByName(name, param, ok);
    if (!ok) return VXHIP_ERR_INVALID;
    *out = vxBuildScene(s);
    return VXHIP_OK;
}

uint32_t vxhip_scene_body_count(const vxhip_scene* s) { return s ? (uint32_t)s->bodies.size() : 0; }

const vx_body* vxhip_scene_bodies(const vxhip_scene* s) { return s ? s->bodies.data() : nullptr; }

const uint32_t* vxhip_scene_worlds(const vxhip_scene* s, uint32_t* worldCount) {
    if (worldCount) *worldCount = s ? s->worlds : 0;
    return s ? s->worldOf.data() : nullptr;
}

const vx_joint* vxhip_scene_joints(const vxhip_scene* s, uint32_t* count) {
    if (count) *count = s ? (uint32_t)s->joints.size() : 0;
    return s && !s->joints.empty() ? s->joints.data() : nullptr;
}

int vxhip_scene_soup(const vxhip_scene* s, vx_soup* out) {
    if (!s || !out) return VXHIP_ERR_INVALID;
    out->vertices = s->vertices.data(); out->vertexCount = (uint32_t)s->vertices.size();
    out->indices = s->indices.data(); out->indexCount = (uint32_t)s->indices.size();
    out->meshes = s->meshes.data(); out->meshCount = (uint32_t)s->meshes.size();
    out->bvhNodes = s->nodes.data(); out->bvhNodeCount = (uint32_t)s->nodes.size();
    out->bvhTriRefs = s->triRefs.data(); out->bvhTriRefCount = (uint32_t)s->triRefs.size();
    out->hullPoints = s->hullPoints.data(); out->hullPointCount = (uint32_t)s->hullPoints.size();
    out->hullFaceIndices = s->hullFaces.data(); out->hullFaceIndexCount = (uint32_t)s->hullFaces.size();
    out->compoundChildren = nullptr; out->compoundChildCount = 0;
    return VXHIP_OK;
}

int vxhip_scene_config(const vxhip_scene* s, vxhip_config* cfg, float* dt) {
    if (!s || !cfg) return VXHIP_ERR_INVALID;
    std::memcpy(cfg->gravity, s->gravity, sizeof(cfg->gravity));
    cfg->substeps = s->substeps;
    cfg->options = {1e-5f, 1000.0f, 10.0f, 8, 3, VX_FRICTION_TWO_AXIS, VX_ITERATION_FIXED, 0, 0};
    if (dt) *dt = s->dt;
    return VXHIP_OK;
}

void vxhip_scene_free(vxhip_scene* s) { delete s; }

}  // extern "C"
