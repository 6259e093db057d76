// Deterministic synthetic scene descriptions for the device-step workloads.
//
// A description lists bodies by shape, size, mass and initial state; it holds
// no solver records.  Two builders turn it into bodies: the product's
// scene_build.cpp (vx_body records + flat geometry for the HIP step) and any
// host that owns a Velox World (World::addBox/addSphere/... calls).
//
// Random numbers come from the 32-bit LCG s = 1664525 s + 1013904223 with
// u = (s >> 8) / 2^24, the generator Velox's own fuzz scenes use
// (examples/fuzz_demo.cpp:23-31).  Orientations are uniform (Shoemake's
// subgroup algorithm); normals use Box-Muller.
#pragma once
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <vector>

namespace vxscene {

struct Rng {
    uint32_t s;
    explicit Rng(uint32_t seed) : s(seed ? seed : 1u) {}
    float next() { s = s * 1664525u + 1013904223u; return (s >> 8) * (1.0f / 16777216.0f); }
    float range(float lo, float hi) { return lo + (hi - lo) * next(); }
    int pick(int n) { return (int)(next() * n) % n; }
    float normal() {  // Box-Muller, one value per call
        float u1 = next(), u2 = next();
        if (u1 < 1e-7f) u1 = 1e-7f;
        return sqrtf(-2.0f * logf(u1)) * cosf(6.2831853f * u2);
    }
    void quat(float q[4]) {  // Shoemake: uniform on SO(3)
        float u1 = next(), u2 = next(), u3 = next();
        float a = sqrtf(1.0f - u1), b = sqrtf(u1);
        q[0] = a * sinf(6.2831853f * u2);
        q[1] = a * cosf(6.2831853f * u2);
        q[2] = b * sinf(6.2831853f * u3);
        q[3] = b * cosf(6.2831853f * u3);
    }
};

enum Kind : uint8_t { Sphere, Box, Capsule, Hull, Plane, Heightfield };

struct BodySpec {
    Kind kind = Sphere;
    float mass = 1.0f;            // 0 => static
    float radius = 0.5f;          // sphere / capsule
    float half[3] = {0.5f, 0.5f, 0.5f};
    float halfHeight = 0.5f;      // capsule
    float pos[3] = {0, 0, 0};
    float quat[4] = {0, 0, 0, 1};
    float vel[3] = {0, 0, 0};
    float angVel[3] = {0, 0, 0};
    float friction = 0.5f;        // Body default (body.h:143)
    float restitution = 0.3f;     // Body default (body.h:142)
    float normal[3] = {0, 1, 0};  // plane
    float offset = 0.0f;          // plane
    uint32_t hullFirst = 0, hullCount = 0;  // into Scene::hullPoints (xyz triples)
    uint32_t hfIndex = 0;         // into Scene::heightfields
    uint32_t world = 0;           // batched scenes: owning world
    uint8_t enableSleep = 0;      // Body::enableSleep (body.h:155)
    uint8_t quality = 1;          // BodyCcdTuning::quality (ccd.h:13): 1 Medium, 2 High
};

// A ball joint between bodies a and b (indices into Scene::bodies) at local
// anchors; every other Joint field keeps the reference's default.
struct JointSpec {
    uint32_t a = 0, b = 0;
    float anchorA[3] = {0, 0, 0}, anchorB[3] = {0, 0, 0};
};

struct HeightfieldSpec {
    uint32_t width = 2, depth = 2;
    float cell = 1.0f;
    float origin[3] = {0, 0, 0};
    std::vector<float> heights;
};

struct Scene {
    std::vector<BodySpec> bodies;
    std::vector<JointSpec> joints;
    std::vector<float> hullPoints;      // xyz triples
    std::vector<HeightfieldSpec> heightfields;
    float gravity[3] = {0, -9.81f, 0};
    float dt = 1.0f / 60.0f;
    int substeps = 4;
    int velocityIterations = 8;
    uint32_t worlds = 1;
};

inline BodySpec plane(float friction = 0.5f, float restitution = 0.3f) {
    BodySpec b; b.kind = Plane; b.mass = 0.0f; b.friction = friction; b.restitution = restitution; return b;
}

// Volume of the convex hull of `n` points in general position (brute-force
// facet enumeration; n is small).  Used only to turn a density into a mass.
inline double hullVolume(const float* p, uint32_t n) {
    double cx = 0, cy = 0, cz = 0;
    for (uint32_t i = 0; i < n; ++i) { cx += p[3 * i]; cy += p[3 * i + 1]; cz += p[3 * i + 2]; }
    cx /= n; cy /= n; cz /= n;
    double vol = 0.0;
    for (uint32_t i = 0; i < n; ++i)
        for (uint32_t j = i + 1; j < n; ++j)
            for (uint32_t k = j + 1; k < n; ++k) {
                double ax = p[3 * i], ay = p[3 * i + 1], az = p[3 * i + 2];
                double ux = p[3 * j] - ax, uy = p[3 * j + 1] - ay, uz = p[3 * j + 2] - az;
                double vx = p[3 * k] - ax, vy = p[3 * k + 1] - ay, vz = p[3 * k + 2] - az;
                double nx = uy * vz - uz * vy, ny = uz * vx - ux * vz, nz = ux * vy - uy * vx;
                int pos = 0, neg = 0;
                for (uint32_t m = 0; m < n && !(pos && neg); ++m) {
                    if (m == i || m == j || m == k) continue;
                    double d = nx * (p[3 * m] - ax) + ny * (p[3 * m + 1] - ay) + nz * (p[3 * m + 2] - az);
                    if (d > 1e-12) ++pos; else if (d < -1e-12) ++neg;
                }
                if (pos && neg) continue;
                double side = nx * (cx - ax) + ny * (cy - ay) + nz * (cz - az);
                vol += fabs(side) / 6.0;  // tetra (centroid, facet)
            }
    return vol;
}

// Velox's demo scenes seed the same way; one seed per workload.
constexpr uint32_t kSeedLattice = 1001u, kSeedDebris = 2002u, kSeedCcd = 3003u,
                   kSeedBatch = 4004u, kSeedDiff = 5005u,
                   kSeedSleep = 6006u, kSeedToi = 7007u, kSeedSpheres = 8008u;

// 1,000 boxes (half-extent 0.5, density 1000) in a 10x10x10 lattice at 1.1 m
// spacing with U(-0.05,0.05) jitter and uniform orientations, on a static
// plane; friction 0.6, restitution 0.
inline Scene lattice(int nx = 10, int ny = 10, int nz = 10, uint32_t seed = kSeedLattice) {
    Scene s;
    Rng r(seed);
    s.bodies.push_back(plane(0.6f, 0.0f));
    for (int i = 0; i < nx; ++i)
        for (int j = 0; j < ny; ++j)
            for (int k = 0; k < nz; ++k) {
                BodySpec b;
                b.kind = Box;
                b.mass = 1000.0f;
                b.pos[0] = 1.1f * i - 4.95f + r.range(-0.05f, 0.05f);
                b.pos[1] = 1.1f * j + 1.0f + r.range(-0.05f, 0.05f);
                b.pos[2] = 1.1f * k - 4.95f + r.range(-0.05f, 0.05f);
                r.quat(b.quat);
                b.friction = 0.6f;
                b.restitution = 0.0f;
                s.bodies.push_back(b);
            }
    return s;
}

// Mixed debris: 30% spheres r~U(.1,.5), 30% boxes h~U(.1,.5)^3, 20% capsules
// r~U(.05,.2) hh~U(.1,.4), 20% hulls of 24 points on an ellipsoid with semi-axes
// U(.2,.6); positions uniform in [-X/2,X/2]x[1,1+Y]x[-Z/2,Z/2], v~N(0,2),
// density 1000, above one static plane.  The default is the 50,000-body case.
inline Scene debris(uint32_t count = 50000, float X = 100.0f, float Y = 50.0f, float Z = 100.0f,
                    uint32_t seed = kSeedDebris) {
    Scene s;
    Rng r(seed);
    s.bodies.push_back(plane());
    std::vector<uint8_t> kinds(count);
    uint32_t nS = count * 3 / 10, nB = count * 3 / 10, nC = count / 5;
    for (uint32_t i = 0; i < count; ++i)
        kinds[i] = i < nS ? Sphere : i < nS + nB ? Box : i < nS + nB + nC ? Capsule : Hull;
    for (uint32_t i = count; i > 1; --i) {  // Fisher-Yates
        uint32_t j = (uint32_t)r.pick((int)i);
        uint8_t t = kinds[i - 1]; kinds[i - 1] = kinds[j]; kinds[j] = t;
    }
    const float rho = 1000.0f, pi = 3.14159265f;
    for (uint32_t i = 0; i < count; ++i) {
        BodySpec b;
        b.kind = (Kind)kinds[i];
        switch (b.kind) {
        case Sphere:
            b.radius = r.range(0.1f, 0.5f);
            b.mass = rho * (4.0f / 3.0f) * pi * b.radius * b.radius * b.radius;
            break;
        case Box:
            for (int a = 0; a < 3; ++a) b.half[a] = r.range(0.1f, 0.5f);
            b.mass = rho * 8.0f * b.half[0] * b.half[1] * b.half[2];
            break;
        case Capsule:
            b.radius = r.range(0.05f, 0.2f);
            b.halfHeight = r.range(0.1f, 0.4f);
            b.mass = rho * (pi * b.radius * b.radius * 2.0f * b.halfHeight +
                            (4.0f / 3.0f) * pi * b.radius * b.radius * b.radius);
            break;
        default: {
            float ax[3] = {r.range(0.2f, 0.6f), r.range(0.2f, 0.6f), r.range(0.2f, 0.6f)};
            b.hullFirst = (uint32_t)(s.hullPoints.size() / 3);
            b.hullCount = 24;
            for (int k = 0; k < 24; ++k) {
                float d[3] = {r.normal(), r.normal(), r.normal()};
                float len = sqrtf(d[0] * d[0] + d[1] * d[1] + d[2] * d[2]);
                if (len < 1e-6f) { d[0] = 1; d[1] = 0; d[2] = 0; len = 1; }
                for (int a = 0; a < 3; ++a) s.hullPoints.push_back(ax[a] * d[a] / len);
            }
            b.mass = rho * (float)hullVolume(&s.hullPoints[3 * b.hullFirst], 24);
            break;
        }
        }
        b.pos[0] = r.range(-0.5f * X, 0.5f * X);
        b.pos[1] = r.range(1.0f, 1.0f + Y);
        b.pos[2] = r.range(-0.5f * Z, 0.5f * Z);
        r.quat(b.quat);
        for (int a = 0; a < 3; ++a) b.vel[a] = 2.0f * r.normal();
        s.bodies.push_back(b);
    }
    return s;
}

// Contact-free flight: bodies far apart (no plane), spinning, for integration
// parity over many steps.
inline Scene freeFlight(uint32_t count = 128, uint32_t seed = 77u) {
    Scene s;
    Rng r(seed);
    for (uint32_t i = 0; i < count; ++i) {
        BodySpec b;
        b.kind = (Kind)(i % 4);
        b.radius = r.range(0.2f, 0.5f);
        b.halfHeight = r.range(0.2f, 0.6f);
        for (int a = 0; a < 3; ++a) b.half[a] = r.range(0.1f, 0.9f);
        if (b.kind == Hull) {
            b.hullFirst = (uint32_t)(s.hullPoints.size() / 3);
            b.hullCount = 12;
            for (int k = 0; k < 12; ++k) {
                float d[3] = {r.normal(), r.normal(), r.normal()};
                float len = sqrtf(d[0] * d[0] + d[1] * d[1] + d[2] * d[2]);
                for (int a = 0; a < 3; ++a) s.hullPoints.push_back(b.half[a] * d[a] / len);
            }
        }
        b.mass = r.range(0.5f, 20.0f);
        b.pos[0] = 40.0f * (float)(i % 8);
        b.pos[1] = 40.0f * (float)((i / 8) % 8) + 500.0f;
        b.pos[2] = 40.0f * (float)(i / 64);
        r.quat(b.quat);
        for (int a = 0; a < 3; ++a) { b.vel[a] = r.range(-3, 3); b.angVel[a] = r.range(-6, 6); }
        s.bodies.push_back(b);
    }
    return s;
}

// Dense mixed pile for parity tests: a few hundred bodies dropped into a
// small volume so that manifolds, EPA and deep stacks all occur.
inline Scene pile(uint32_t count = 240, float X = 6.0f, float Y = 8.0f, uint32_t seed = 606u) {
    Scene s = debris(count, X, Y, X, seed);
    return s;
}

// Sphere pairs at every separation the sphere-sphere contact distinguishes:
// `count` spheres r~U(.1,.5) at uniform positions in a `side`-metre cube
// (overlapping, touching and separated neighbours), v~N(0,1) so the
// speculative reach varies per pair, plus eight pairs of spheres whose centres
// coincide exactly and, for each of 64 seeded spheres, a partner placed at a
// distance within a few float steps of the pair's speculative reach.  No plane.
inline Scene sphereCluster(uint32_t count = 600, float side = 5.0f, uint32_t seed = kSeedSpheres) {
    Scene s;
    Rng r(seed);
    auto sphere = [&](float x, float y, float z, float rad) {
        BodySpec b;
        b.kind = Sphere;
        b.radius = rad;
        b.mass = 1000.0f * (4.0f / 3.0f) * 3.14159265f * rad * rad * rad;
        b.pos[0] = x; b.pos[1] = y; b.pos[2] = z;
        for (int a = 0; a < 3; ++a) b.vel[a] = r.normal();
        s.bodies.push_back(b);
        return s.bodies.size() - 1;
    };
    for (uint32_t i = 0; i < count; ++i)
        sphere(r.range(0.0f, side), r.range(0.0f, side), r.range(0.0f, side), r.range(0.1f, 0.5f));
    for (int k = 0; k < 8; ++k) {  // coincident centres
        const float x = side + 2.0f + 3.0f * (float)k, y = 1.0f, z = 1.0f;
        sphere(x, y, z, r.range(0.1f, 0.5f));
        sphere(x, y, z, r.range(0.1f, 0.5f));
    }
    for (int k = 0; k < 64; ++k) {  // partners at the edge of the speculative reach
        const float x = side + 2.0f + 3.0f * (float)(k % 8), y = 4.0f + 3.0f * (float)(k / 8), z = 1.0f;
        const size_t a = sphere(x, y, z, r.range(0.1f, 0.5f));
        const float ra = s.bodies[a].radius, rb = r.range(0.1f, 0.5f);
        // maxPointSpeed of a sphere is its speed (no angular velocity); the
        // reach is ra + rb + (|va| + |vb|) * dt + 1e-3, placed along x.
        BodySpec b;
        const float va = sqrtf(s.bodies[a].vel[0] * s.bodies[a].vel[0] + s.bodies[a].vel[1] * s.bodies[a].vel[1] +
                               s.bodies[a].vel[2] * s.bodies[a].vel[2]);
        float vb[3] = {r.normal(), r.normal(), r.normal()};
        const float vbl = sqrtf(vb[0] * vb[0] + vb[1] * vb[1] + vb[2] * vb[2]);
        const float reach = ra + rb + (va + vbl) * s.dt + 1e-3f;
        const float off = reach * (1.0f + 1e-5f * (float)((int)(k % 9) - 4));
        b.kind = Sphere;
        b.radius = rb;
        b.mass = 1000.0f * (4.0f / 3.0f) * 3.14159265f * rb * rb * rb;
        b.pos[0] = x + off; b.pos[1] = y; b.pos[2] = z;
        b.vel[0] = vb[0]; b.vel[1] = vb[1]; b.vel[2] = vb[2];
        s.bodies.push_back(b);
    }
    return s;
}

// Axis-aligned stack of boxes (half-extent `h`, faces 0.5 mm apart, so every
// neighbour pair is inside the 1 mm speculative slop at rest) with its
// lowest layer 0.5 mm above the plane, centred at (cx, cz).
inline void boxStack(Scene& s, int nx, int ny, int nz, float h, float cx, float cz, float mass) {
    const float gap = 5e-4f, pitch = 2.0f * h + gap;
    for (int i = 0; i < nx; ++i)
        for (int j = 0; j < ny; ++j)
            for (int k = 0; k < nz; ++k) {
                BodySpec b;
                b.kind = Box;
                b.half[0] = b.half[1] = b.half[2] = h;
                b.mass = mass;
                b.pos[0] = cx + pitch * ((float)i - 0.5f * (float)(nx - 1));
                b.pos[1] = h + gap + pitch * (float)j;
                b.pos[2] = cz + pitch * ((float)k - 0.5f * (float)(nz - 1));
                b.friction = 0.6f;
                b.restitution = 0.0f;
                s.bodies.push_back(b);
            }
}

// The crowded regime of a large settled pile, reached at step 1: one island
// of 55,000-90,000 contacts (a 24 x 5 x 24 stack of 0.5 m boxes), a 6 m slab
// resting on its top layer with a 10 x 10 grid of spheres on the slab (the
// slab touches ~680 contacts, so the greedy colouring runs to ~700 colours,
// far past the 64 colours of one mask word), a separate 8 x 3 x 8 stack of
// about 3,300 contacts, and 64 lone spheres on the plane.  Friction 0.6,
// restitution 0.  `small`: a 12 x 1 x 12 layer under the slab and spheres
// only (the slab still takes ~680 colours), for per-stage snapshots.
inline Scene slabStack(bool small = false) {
    Scene s;
    s.bodies.push_back(plane(0.6f, 0.0f));
    const float h = 0.25f, gap = 5e-4f, pitch = 2.0f * h + gap;
    const int layers = small ? 1 : 5;
    boxStack(s, small ? 12 : 24, layers, small ? 12 : 24, h, 0.0f, 0.0f, 125.0f);
    const float top = gap + (float)layers * pitch - gap;  // upper face of the top layer
    BodySpec slab;
    slab.kind = Box;
    slab.half[0] = 3.0f; slab.half[1] = 0.15f; slab.half[2] = 3.0f;
    slab.mass = 500.0f;
    slab.pos[1] = top + gap + slab.half[1];
    slab.friction = 0.6f;
    slab.restitution = 0.0f;
    s.bodies.push_back(slab);
    const float slabTop = slab.pos[1] + slab.half[1];
    for (int i = 0; i < 10; ++i)
        for (int k = 0; k < 10; ++k) {
            BodySpec b;
            b.kind = Sphere;
            b.radius = 0.2f;
            b.mass = 33.5f;
            b.pos[0] = 0.55f * ((float)i - 4.5f);
            b.pos[1] = slabTop + gap + b.radius;
            b.pos[2] = 0.55f * ((float)k - 4.5f);
            b.friction = 0.6f;
            b.restitution = 0.0f;
            s.bodies.push_back(b);
        }
    if (small) return s;
    boxStack(s, 8, 3, 8, h, 20.0f, 0.0f, 125.0f);
    for (int i = 0; i < 64; ++i) {
        BodySpec b;
        b.kind = Sphere;
        b.radius = 0.3f;
        b.mass = 113.0f;
        b.pos[0] = -20.0f + 1.0f * (float)(i % 8);
        b.pos[1] = b.radius + gap;
        b.pos[2] = 1.0f * (float)(i / 8) - 3.5f;
        b.friction = 0.6f;
        b.restitution = 0.0f;
        s.bodies.push_back(b);
    }
    return s;
}

// Static heightfield terrain (value-noise octaves) with mixed bodies above it.
inline Scene terrain(uint32_t side = 24, float cell = 0.5f, float amp = 1.0f, uint32_t count = 96,
                     uint32_t seed = 909u) {
    Scene s;
    Rng r(seed);
    HeightfieldSpec hf;
    hf.width = hf.depth = side;
    hf.cell = cell;
    hf.origin[0] = -0.5f * cell * (side - 1);
    hf.origin[1] = 0.0f;
    hf.origin[2] = -0.5f * cell * (side - 1);
    hf.heights.resize((size_t)side * side);
    for (uint32_t z = 0; z < side; ++z)
        for (uint32_t x = 0; x < side; ++x)
            hf.heights[(size_t)z * side + x] = amp * (0.5f * sinf(0.37f * x) * cosf(0.29f * z) + 0.25f * sinf(1.3f * x + 0.7f * z));
    s.heightfields.push_back(hf);
    BodySpec m; m.kind = Heightfield; m.mass = 0.0f; m.hfIndex = 0;
    s.bodies.push_back(m);
    float half = 0.5f * cell * (side - 1) - 1.0f;
    Scene d = debris(count, 2.0f * half, 4.0f, 2.0f * half, seed + 1);
    s.hullPoints = d.hullPoints;
    for (size_t i = 1; i < d.bodies.size(); ++i) {
        BodySpec b = d.bodies[i];
        b.pos[1] += 1.5f * amp;
        s.bodies.push_back(b);
    }
    return s;
}

// Batched worlds: `worlds` independent worlds, each on its own static plane
// (all at the same coordinates; a world id separates them).  A world holds
// n_w = round(exp(U(ln 8, ln 512))) dynamic bodies built from seeded
// structures on a 4 m site grid: box towers of height U{3..12} (unit boxes,
// small yaw jitter), sphere piles (r ~ U(0.2, 0.4), a loose cluster dropped
// from 1-3 m) and capsule chains (3-5 capsules, r ~ U(0.1, 0.15), half-height
// U(0.15, 0.3), lying end to end along x 2 cm apart, each joined to the next
// by a ball joint midway between them).  Density 1000, friction 0.6,
// restitution 0.
inline Scene batchWorlds(uint32_t worlds = 4096, uint32_t seed = kSeedBatch) {
    Scene s;
    s.worlds = worlds;
    Rng r(seed);
    const float pi = 3.14159265f;
    for (uint32_t w = 0; w < worlds; ++w) {
        BodySpec pl = plane(0.6f, 0.0f);
        pl.world = w;
        s.bodies.push_back(pl);
        const uint32_t target = (uint32_t)std::lround(std::exp(r.range(std::log(8.0f), std::log(512.0f))));
        uint32_t placed = 0, site = 0;
        while (placed < target) {
            const float sx = 4.0f * (float)(site % 8) - 14.0f, sz = 4.0f * (float)(site / 8) - 14.0f;
            ++site;
            const uint32_t left = target - placed;
            const int kind = r.pick(3);
            auto spec = [&](Kind k) {
                BodySpec b;
                b.kind = k;
                b.friction = 0.6f;
                b.restitution = 0.0f;
                b.world = w;
                return b;
            };
            if (kind == 0) {  // box tower
                const uint32_t h = std::min<uint32_t>(left, 3u + (uint32_t)r.pick(10));
                for (uint32_t k = 0; k < h; ++k) {
                    BodySpec b = spec(Box);
                    b.half[0] = b.half[1] = b.half[2] = 0.5f;
                    b.mass = 1000.0f;
                    b.pos[0] = sx + r.range(-0.02f, 0.02f);
                    b.pos[1] = 0.5f + 1.002f * (float)k;
                    b.pos[2] = sz + r.range(-0.02f, 0.02f);
                    const float yaw = r.range(-0.1f, 0.1f);
                    b.quat[0] = 0.0f; b.quat[1] = std::sin(0.5f * yaw); b.quat[2] = 0.0f; b.quat[3] = std::cos(0.5f * yaw);
                    s.bodies.push_back(b);
                }
                placed += h;
            } else if (kind == 1) {  // sphere pile
                const uint32_t m = std::min<uint32_t>(left, 8u + (uint32_t)r.pick(25));
                for (uint32_t k = 0; k < m; ++k) {
                    BodySpec b = spec(Sphere);
                    b.radius = r.range(0.2f, 0.4f);
                    b.mass = 1000.0f * (4.0f / 3.0f) * pi * b.radius * b.radius * b.radius;
                    b.pos[0] = sx + 0.85f * (float)(k % 3) - 0.85f + r.range(-0.05f, 0.05f);
                    b.pos[1] = 1.0f + 0.85f * (float)(k / 9) + r.range(0.0f, 0.1f);
                    b.pos[2] = sz + 0.85f * (float)((k / 3) % 3) - 0.85f + r.range(-0.05f, 0.05f);
                    s.bodies.push_back(b);
                }
                placed += m;
            } else {  // capsule chain, axes along x, end to end
                const uint32_t m = std::min<uint32_t>(left, 3u + (uint32_t)r.pick(3));
                const float rad = r.range(0.1f, 0.15f), hh = r.range(0.15f, 0.3f);
                const float link = 2.0f * (hh + rad) + 0.02f;  // centre spacing
                const uint32_t first = (uint32_t)s.bodies.size();
                for (uint32_t k = 0; k < m; ++k) {
                    BodySpec b = spec(Capsule);
                    b.radius = rad;
                    b.halfHeight = hh;
                    b.mass = 1000.0f * (pi * rad * rad * 2.0f * hh + (4.0f / 3.0f) * pi * rad * rad * rad);
                    b.pos[0] = sx + link * ((float)k - 0.5f * (float)(m - 1));
                    b.pos[1] = rad + 0.01f;
                    b.pos[2] = sz;
                    // +90 degrees about z: the capsule's local y axis along world -x.
                    b.quat[0] = 0.0f; b.quat[1] = 0.0f; b.quat[2] = std::sin(0.25f * pi); b.quat[3] = std::cos(0.25f * pi);
                    s.bodies.push_back(b);
                }
                for (uint32_t k = 0; k + 1 < m; ++k) {
                    JointSpec j;
                    j.a = first + k;
                    j.b = first + k + 1;
                    j.anchorA[1] = -0.5f * link;  // world +x of link k
                    j.anchorB[1] = 0.5f * link;   // world -x of link k + 1
                    s.joints.push_back(j);
                }
                placed += m;
            }
        }
    }
    return s;
}

// Differentiable worlds: `worlds` worlds, each a static plane (friction 0.6)
// and 64 cubes (half-extent 0.2 m) in an 8 x 8 grid at 0.6 m pitch resting
// 0.5 mm above it, with mass U(0.5, 5) kg, friction U(0.2, 1.0) (geometric-mean
// combine) and restitution 0, each pushed by an initial impulse: linear
// velocity (U(0.5, 1.5), 0, U(-0.25, 0.25)) m/s and a spin about the vertical
// of U(-0.5, 0.5) rad/s.  The cubes slide and stop within about 0.35 m, so
// neighbours meet gently if at all (no pair comes near the CCD net's
// penetration threshold).  Cubes have isotropic inertia, which the adjoint
// assumes.
inline Scene diffWorlds(uint32_t worlds = 1024, uint32_t boxes = 64, uint32_t seed = kSeedDiff) {
    Scene s;
    s.worlds = worlds;
    Rng r(seed);
    const uint32_t side = (uint32_t)std::ceil(std::sqrt((float)boxes));
    for (uint32_t w = 0; w < worlds; ++w) {
        BodySpec pl = plane(0.6f, 0.0f);
        pl.world = w;
        s.bodies.push_back(pl);
        for (uint32_t k = 0; k < boxes; ++k) {
            BodySpec b;
            b.kind = Box;
            b.world = w;
            b.half[0] = b.half[1] = b.half[2] = 0.2f;
            b.mass = r.range(0.5f, 5.0f);
            b.friction = r.range(0.2f, 1.0f);
            b.restitution = 0.0f;
            b.pos[0] = 0.6f * ((float)(k % side) - 0.5f * (float)(side - 1));
            b.pos[1] = 0.2f + 5e-4f;
            b.pos[2] = 0.6f * ((float)(k / side) - 0.5f * (float)(side - 1));
            b.vel[0] = r.range(0.5f, 1.5f);
            b.vel[2] = r.range(-0.25f, 0.25f);
            b.angVel[1] = r.range(-0.5f, 0.5f);
            s.bodies.push_back(b);
        }
    }
    return s;
}

// Sleep and wake: eight towers of three unit boxes (1 mm gaps, U(-0.02,0.02)
// x/z jitter) and eight resting spheres (r = 0.4) on a plane settle and fall
// asleep island by island; the middle box of tower 7 has sleeping disabled, so
// its island stays awake.  A sphere (r = 0.5, 50 kg) released at y = 80 above
// tower 0 strikes it after about 3.9 s and wakes it.  Friction 0.6,
// restitution 0.
inline Scene settle(uint32_t seed = kSeedSleep) {
    Scene s;
    Rng r(seed);
    s.bodies.push_back(plane(0.6f, 0.0f));
    for (int t = 0; t < 8; ++t)
        for (int j = 0; j < 3; ++j) {
            BodySpec b;
            b.kind = Box;
            b.mass = 100.0f;
            b.pos[0] = 3.0f * (float)(t % 4) - 4.5f + r.range(-0.02f, 0.02f);
            b.pos[1] = 0.501f + 1.001f * (float)j;
            b.pos[2] = 3.0f * (float)(t / 4) - 1.5f + r.range(-0.02f, 0.02f);
            b.friction = 0.6f;
            b.restitution = 0.0f;
            b.enableSleep = (t == 7 && j == 1) ? 0 : 1;
            s.bodies.push_back(b);
        }
    for (int k = 0; k < 8; ++k) {
        BodySpec b;
        b.kind = Sphere;
        b.radius = 0.4f;
        b.mass = 20.0f;
        b.pos[0] = 2.0f * (float)k - 7.0f;
        b.pos[1] = 0.401f;
        b.pos[2] = 4.5f;
        b.friction = 0.6f;
        b.restitution = 0.0f;
        b.enableSleep = 1;
        s.bodies.push_back(b);
    }
    BodySpec striker;
    striker.kind = Sphere;
    striker.radius = 0.5f;
    striker.mass = 50.0f;
    striker.pos[0] = s.bodies[1].pos[0];
    striker.pos[1] = 80.0f;
    striker.pos[2] = s.bodies[1].pos[2];
    striker.friction = 0.6f;
    striker.restitution = 0.0f;
    striker.enableSleep = 1;
    s.bodies.push_back(striker);
    return s;
}

// High-quality (MotionQuality::High) bodies for the multi-TOI scheduler: 24
// spheres (r = 0.1 m, 1 kg) fired along +x at 60-240 m/s into a wall of 12
// static boxes, 4 High boxes fired at each other in two pairs at 40 m/s, and
// 16 Medium boxes dropped on the plane away from them.  Friction 0.5,
// restitution 0.3 (Body defaults).
inline Scene highToi(uint32_t seed = kSeedToi) {
    Scene s;
    Rng r(seed);
    s.bodies.push_back(plane());
    for (int j = 0; j < 3; ++j)
        for (int k = 0; k < 4; ++k) {
            BodySpec w;
            w.kind = Box;
            w.mass = 0.0f;
            w.half[0] = 0.25f; w.half[1] = 1.0f; w.half[2] = 1.0f;
            w.pos[0] = 12.0f; w.pos[1] = 1.0f + 2.0f * (float)j; w.pos[2] = -3.0f + 2.0f * (float)k;
            s.bodies.push_back(w);
        }
    for (int i = 0; i < 24; ++i) {
        BodySpec b;
        b.kind = Sphere;
        b.radius = 0.1f;
        b.mass = 1.0f;
        b.pos[0] = r.range(-2.0f, 2.0f);
        b.pos[1] = r.range(0.5f, 5.5f);
        b.pos[2] = r.range(-3.5f, 3.5f);
        b.vel[0] = r.range(60.0f, 240.0f);
        b.vel[1] = r.range(-2.0f, 2.0f);
        b.vel[2] = r.range(-2.0f, 2.0f);
        b.quality = 2;
        s.bodies.push_back(b);
    }
    for (int p = 0; p < 2; ++p)
        for (int side = 0; side < 2; ++side) {
            BodySpec b;
            b.kind = Box;
            b.half[0] = b.half[1] = b.half[2] = 0.2f;
            b.mass = 5.0f;
            b.pos[0] = side ? -4.0f : -12.0f;
            b.pos[1] = 2.0f + 3.0f * (float)p;
            b.pos[2] = 8.0f + 0.05f * (float)side;
            r.quat(b.quat);
            b.vel[0] = side ? -40.0f : 40.0f;
            b.quality = 2;
            s.bodies.push_back(b);
        }
    for (int i = 0; i < 16; ++i) {
        BodySpec b;
        b.kind = Box;
        b.half[0] = b.half[1] = b.half[2] = 0.3f;
        b.mass = 10.0f;
        b.pos[0] = -20.0f + 1.5f * (float)(i % 4);
        b.pos[1] = 0.6f + 1.0f * (float)(i / 4);
        b.pos[2] = -20.0f;
        r.quat(b.quat);
        s.bodies.push_back(b);
    }
    return s;
}

// CCD stress: 4,096 spheres (r = 0.05 m, density 1000) fired at 2 km/s in a
// 10-degree cone at a wall of 100 x 100 stacked unit boxes; 256 unit boxes
// spinning at 1.5 km/s tip speed (1732 rad/s about random axes) above a
// 256 x 256 heightfield at 0.5 m spacing whose heights are five octaves of
// seeded sinusoidal noise scaled to 3 m amplitude.  The reduced copy
// (ccdStressSmall) keeps every ingredient at a size the strict CPU reference
// steps in seconds: a 10 x 10 wall at x = 10.5, 16 x 16 bullets, 4 x 4
// spinning boxes and a 64 x 64 heightfield.
inline Scene ccdStress(uint32_t seed = kSeedCcd, int wall = 100, int bulletSide = 64, int spinSide = 16,
                       uint32_t side = 256, float wallX = 20.5f) {
    Scene s;
    Rng r(seed);
    const float pi = 3.14159265f;
    HeightfieldSpec hf;
    hf.width = hf.depth = side;
    hf.cell = 0.5f;
    hf.origin[0] = hf.origin[2] = -0.5f * hf.cell * (float)(side - 1);
    hf.origin[1] = 0.0f;
    hf.heights.assign((size_t)side * side, 0.0f);
    float phase[5][2], freq = 0.05f, amp = 1.0f, total = 0.0f;
    for (int o = 0; o < 5; ++o) { phase[o][0] = r.range(0.0f, 2.0f * pi); phase[o][1] = r.range(0.0f, 2.0f * pi); }
    for (int o = 0; o < 5; ++o, freq *= 2.0f, amp *= 0.5f) {
        total += amp;
        for (uint32_t z = 0; z < side; ++z)
            for (uint32_t x = 0; x < side; ++x) {
                const float wx = hf.origin[0] + hf.cell * (float)x, wz = hf.origin[2] + hf.cell * (float)z;
                hf.heights[(size_t)z * side + x] += amp * std::sin(freq * wx + phase[o][0]) * std::cos(freq * wz + phase[o][1]);
            }
    }
    for (float& h : hf.heights) h *= 3.0f / total;
    s.heightfields.push_back(hf);
    BodySpec terrain;
    terrain.kind = Heightfield;
    terrain.mass = 0.0f;
    terrain.friction = 0.6f;
    terrain.restitution = 0.0f;
    s.bodies.push_back(terrain);
    auto box = [&](float x, float y, float z) {
        BodySpec b;
        b.kind = Box;
        b.half[0] = b.half[1] = b.half[2] = 0.5f;
        b.mass = 1000.0f;
        b.pos[0] = x; b.pos[1] = y; b.pos[2] = z;
        b.friction = 0.6f;
        b.restitution = 0.0f;
        return b;
    };
    const float W = (float)wall;
    for (int j = 0; j < wall; ++j)  // the wall, x in [wallX - 0.5, wallX + 0.5]
        for (int k = 0; k < wall; ++k)
            s.bodies.push_back(box(wallX, 4.5f + 1.002f * (float)j, -0.5f * (W - 1.0f) + 1.002f * (float)k));
    const float cone = 10.0f * pi / 180.0f;
    const float nb1 = (float)(bulletSide - 1);
    for (int j = 0; j < bulletSide; ++j)  // bullets
        for (int k = 0; k < bulletSide; ++k) {
            BodySpec b;
            b.kind = Sphere;
            b.radius = 0.05f;
            b.mass = 1000.0f * (4.0f / 3.0f) * pi * 0.05f * 0.05f * 0.05f;
            b.pos[0] = -10.0f;
            b.pos[1] = 0.1f * W + 0.8f * W * (float)j / nb1;
            b.pos[2] = -0.4f * W + 0.8f * W * (float)k / nb1;
            // Uniform direction in the cone about +x.
            const float cosT = 1.0f - r.next() * (1.0f - std::cos(cone)), sinT = std::sqrt(1.0f - cosT * cosT);
            const float phi = r.range(0.0f, 2.0f * pi);
            b.vel[0] = 2000.0f * cosT;
            b.vel[1] = 2000.0f * sinT * std::cos(phi);
            b.vel[2] = 2000.0f * sinT * std::sin(phi);
            b.friction = 0.6f;
            b.restitution = 0.0f;
            s.bodies.push_back(b);
        }
    const float omega = 1500.0f / 0.8660254f;
    const float sx = -50.0f * (float)spinSide / 16.0f, sz = -24.0f * (float)spinSide / 16.0f;
    for (int j = 0; j < spinSide; ++j)  // spinning boxes above the terrain
        for (int k = 0; k < spinSide; ++k) {
            BodySpec b = box(sx + 3.0f * (float)j, 6.0f + r.range(0.0f, 2.0f), sz + 3.0f * (float)k);
            r.quat(b.quat);
            float ax = r.normal(), ay = r.normal(), az = r.normal();
            const float l = std::sqrt(ax * ax + ay * ay + az * az);
            b.angVel[0] = omega * ax / l; b.angVel[1] = omega * ay / l; b.angVel[2] = omega * az / l;
            s.bodies.push_back(b);
        }
    return s;
}

inline Scene ccdStressSmall(uint32_t seed = kSeedCcd) { return ccdStress(seed, 10, 16, 4, 64, 10.5f); }

}  // namespace vxscene
