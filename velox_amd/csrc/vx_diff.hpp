// Differentiable form of the device step: per-world replay kernels that
// record what the adjoint needs (a tape of every row visit), and the hand-
// written adjoint kernels of the substepped contact solve, the transform
// advance and the positional correction (vxd_adjoint.hpp holds the math).
//
// Layout.  A step's contacts are in solve order (colour-major over all
// worlds).  `list` holds the solve positions grouped by world, in solve
// order within a world, with world w's entries at [wStart[w], wStart[w+1]);
// within a world, equal colours are consecutive ("runs").  One workgroup
// takes one world: it walks the world's runs in order (the forward) or in
// reverse (the adjoint), one contact per thread, with a workgroup barrier
// between runs.  A run's contacts touch distinct dynamic bodies, so a run is
// free of write conflicts and its order is the colour order's.
//
// The forward rows are those of the product's solve (kPrepareRows and
// solveRegs, the warm start of warmStartContact), so a replay reproduces the
// step bit for bit.  Tape of a visit: (vA, cN), (wA, c1), (vB, c2), (wB, 0)
// before the row.
#pragma once
#include "vx_solve.hpp"
#include "vxd_adjoint.hpp"

namespace vxk {

constexpr int kDiffThreads = 256;
constexpr uint32_t kDiffMaxRuns = 2048;  // colour runs per world held in LDS

// (world of the contact, solve position) for the world grouping; per-world
// contact counts and the largest world.
__global__ __launch_bounds__(kThreads) void kDiffWorldKeys(const vx_contact* __restrict__ cs, uint32_t m,
                                                           const uint32_t* __restrict__ worldOf,
                                                           uint64_t* __restrict__ keys, uint32_t* __restrict__ vals,
                                                           uint32_t* __restrict__ count) {
    uint32_t k = blockIdx.x * kThreads + threadIdx.x;
    if (k >= m) return;
    const uint32_t w = worldOf[cs[k].a];
    keys[k] = w;
    vals[k] = k;
    atomicAdd(&count[w], 1u);
}

// The runs of one world's list into LDS (run starts, then the end).
__device__ inline uint32_t diffRuns(const uint32_t* __restrict__ list, const uint32_t* __restrict__ colorOf,
                                    uint32_t s0, uint32_t s1, uint32_t* runStart, uint32_t* overflow) {
    __shared__ uint32_t runsS;
    if (threadIdx.x == 0) {
        uint32_t runs = 0;
        for (uint32_t i = s0; i < s1; ++i)
            if (i == s0 || colorOf[list[i]] != colorOf[list[i - 1]]) {
                if (runs < kDiffMaxRuns) runStart[runs] = i;
                ++runs;
            }
        if (runs > kDiffMaxRuns) {
            atomicAdd(overflow, 1u);
            runs = 0;
        }
        runStart[runs] = s1;
        runsS = runs;
    }
    __syncthreads();
    return runsS;
}

__device__ inline vxadj::V<float> av(float4 v) { return {v.x, v.y, v.z}; }
__device__ inline float4 f4(vxadj::V<float> v, float w) { return make_float4(v.x, v.y, v.z, w); }

// The warm start of one contact on the compact arrays (warmStartContact,
// narrowphase.h:80-124), in the product's operation order.
__device__ inline void warmRow(SolveArrays sa, uint32_t p) {
    const uint4 meta = sa.meta[p];
    if (meta.z & kSensorRow) return;
    const float4 i0 = sa.imp[2 * p], i1 = sa.imp[2 * p + 1];
    const float cN = i0.x, cT1 = i0.y, cT2 = i0.z, cR1 = i0.w, cR2 = i1.x, cS = i1.y;
    if (cN <= 0.0f && cT1 == 0.0f && cT2 == 0.0f && cR1 == 0.0f && cR2 == 0.0f && cS == 0.0f) return;
    const float4* row = sa.rows + 8 * (size_t)p;
    const V3 ra = xyz(row[0]), rb = xyz(row[1]), n = xyz(row[2]), t1 = xyz(row[3]), t2 = xyz(row[4]);
    const V3 imp = n * cN + t1 * cT1 + t2 * cT2;
    const V3 ang = t1 * cR1 + t2 * cR2 + n * cS;
    for (int side = 0; side < 2; ++side) {
        const bool dyn = meta.z & (side ? kDynB : kDynA), rot = meta.z & (side ? kRotB : kRotA);
        if (!dyn) continue;
        const uint32_t body = side ? meta.y : meta.x;
        const float4 fr = sa.frame[body], ii = sa.invI[body];
        V3 v = xyz(sa.vel[body]), w = xyz(sa.ang[body]);
        const V3 r = side ? rb : ra;
        const V3 dv = imp * ii.w;
        v = v + (side ? V3{-dv.x, -dv.y, -dv.z} : dv);
        const V3 dw1 = frameMul(fr, ii, cross(r, imp), rot);
        w = w + (side ? V3{-dw1.x, -dw1.y, -dw1.z} : dw1);
        const V3 dw2 = frameMul(fr, ii, ang, rot);
        w = w + (side ? V3{-dw2.x, -dw2.y, -dw2.z} : dw2);
        sa.vel[body] = make_float4(v.x, v.y, v.z, 0.0f);
        sa.ang[body] = make_float4(w.x, w.y, w.z, 0.0f);
    }
}

__device__ inline void tapeVisit(SolveArrays sa, uint32_t p, float4* __restrict__ tape) {
    const uint4 meta = sa.meta[p];
    const float4 i0 = sa.imp[2 * p];
    tape[0] = make_float4(sa.vel[meta.x].x, sa.vel[meta.x].y, sa.vel[meta.x].z, i0.x);
    tape[1] = make_float4(sa.ang[meta.x].x, sa.ang[meta.x].y, sa.ang[meta.x].z, i0.y);
    tape[2] = make_float4(sa.vel[meta.y].x, sa.vel[meta.y].y, sa.vel[meta.y].z, i0.z);
    tape[3] = make_float4(sa.ang[meta.y].x, sa.ang[meta.y].y, sa.ang[meta.y].z, 0.0f);
}

// One substep's velocity solve of every world, replayed in run order with a
// tape: the warm start (when `warm`) as pass `sweeps`, then the sweeps.
__global__ __launch_bounds__(kDiffThreads) void kDiffSweeps(SolveArrays sa, const uint32_t* __restrict__ list,
                                                            const uint32_t* __restrict__ wStart,
                                                            const uint32_t* __restrict__ colorOf, uint32_t m,
                                                            int sweeps, int warm, float4* __restrict__ tape,
                                                            uint32_t* __restrict__ overflow) {
    __shared__ uint32_t runStart[kDiffMaxRuns + 1];
    const uint32_t w = blockIdx.x, s0 = wStart[w], s1 = wStart[w + 1];
    if (s0 == s1) return;
    const uint32_t runs = diffRuns(list, colorOf, s0, s1, runStart, overflow);
    if (warm) {
        for (uint32_t r = 0; r < runs; ++r) {
            for (uint32_t i = runStart[r] + threadIdx.x; i < runStart[r + 1]; i += kDiffThreads) {
                const uint32_t p = list[i];
                tapeVisit(sa, p, tape + 4 * ((size_t)sweeps * m + i));
                warmRow(sa, p);
            }
            __syncthreads();
        }
    }
    for (int it = 0; it < sweeps; ++it)
        for (uint32_t r = 0; r < runs; ++r) {
            for (uint32_t i = runStart[r] + threadIdx.x; i < runStart[r + 1]; i += kDiffThreads) {
                const uint32_t p = list[i];
                tapeVisit(sa, p, tape + 4 * ((size_t)it * m + i));
                solveRow<false>(sa, p);
            }
            __syncthreads();
        }
}

// Row data of solve position p for the adjoint (isotropic inverse inertia:
// invI.x stands for the whole tensor).
__device__ inline void diffRow(SolveArrays sa, uint32_t p, float h, vxadj::Row<float>& R, vxadj::Side<float>& A,
                               vxadj::Side<float>& B, uint4& meta) {
    meta = sa.meta[p];
    const float4* row = sa.rows + 8 * (size_t)p;
    const float4 r0 = row[0], r1 = row[1], r2 = row[2], r3 = row[3], r4 = row[4], r5 = row[5];
    R.ra = av(r0); R.rb = av(r1); R.n = av(r2); R.t1 = av(r3); R.t2 = av(r4);
    R.kN = r0.w; R.k1 = r1.w; R.k2 = r2.w; R.lg = r4.w; R.h = h;
    R.vn0 = r5.x; R.rest = r5.y; R.mu1 = r5.z; R.mu2 = r5.w;
    const float4 ia = sa.invI[meta.x], ib = sa.invI[meta.y];
    A = {ia.w, ia.x, (meta.z & kDynA) != 0};
    B = {ib.w, ib.x, (meta.z & kDynB) != 0};
}

// Per-contact adjoint accumulator: 4 float4 (ra, lg), (rb, kN), (k1, k2, mu1,
// mu2), (imA, imB, iiA, iiB).
__device__ inline void loadBar(const float4* b, vxadj::RowBar<float>& r) {
    r.ra = av(b[0]); r.lg = b[0].w; r.rb = av(b[1]); r.kN = b[1].w;
    r.k1 = b[2].x; r.k2 = b[2].y; r.mu1 = b[2].z; r.mu2 = b[2].w;
    r.imA = b[3].x; r.imB = b[3].y; r.iiA = b[3].z; r.iiB = b[3].w;
}
__device__ inline void storeBar(float4* b, const vxadj::RowBar<float>& r) {
    b[0] = f4(r.ra, r.lg); b[1] = f4(r.rb, r.kN);
    b[2] = make_float4(r.k1, r.k2, r.mu1, r.mu2);
    b[3] = make_float4(r.imA, r.imB, r.iiA, r.iiB);
}

// The adjoint of kDiffSweeps: the sweeps in reverse, then the warm start.
// adjV / adjW: per-body velocity adjoints (in and out); lamBar: per solve
// position (cN, c1, c2) adjoints (in and out); rowBar accumulates.
__global__ __launch_bounds__(kDiffThreads) void kDiffSweepsAdj(SolveArrays sa, const uint32_t* __restrict__ list,
                                                               const uint32_t* __restrict__ wStart,
                                                               const uint32_t* __restrict__ colorOf, uint32_t m,
                                                               int sweeps, int warm, float h,
                                                               const float4* __restrict__ tape,
                                                               float4* __restrict__ adjV, float4* __restrict__ adjW,
                                                               float4* __restrict__ lamBar, float4* __restrict__ rowBar,
                                                               uint32_t* __restrict__ overflow) {
    __shared__ uint32_t runStart[kDiffMaxRuns + 1];
    const uint32_t w = blockIdx.x, s0 = wStart[w], s1 = wStart[w + 1];
    if (s0 == s1) return;
    const uint32_t runs = diffRuns(list, colorOf, s0, s1, runStart, overflow);
    auto visit = [&](uint32_t i, int pass, bool isWarm) {
        const uint32_t p = list[i];
        vxadj::Row<float> R;
        vxadj::Side<float> A, B;
        uint4 meta;
        diffRow(sa, p, h, R, A, B, meta);
        if (meta.z & (isWarm ? kSensorRow : kSkip)) return;
        const float4* t = tape + 4 * ((size_t)pass * m + i);
        const vxadj::Visit<float> pre{av(t[0]), av(t[1]), av(t[2]), av(t[3]), t[0].w, t[1].w, t[2].w};
        const float4 lb = lamBar[p];
        vxadj::Visit<float> bar{av(adjV[meta.x]), av(adjW[meta.x]), av(adjV[meta.y]), av(adjW[meta.y]),
                                lb.x, lb.y, lb.z};
        vxadj::RowBar<float> rb;
        loadBar(rowBar + 4 * (size_t)p, rb);
        if (isWarm) vxadj::warmAdjoint(R, A, B, pre, bar, rb);
        else vxadj::rowAdjoint(R, A, B, pre, bar, rb);
        storeBar(rowBar + 4 * (size_t)p, rb);
        lamBar[p] = make_float4(bar.cN, bar.c1, bar.c2, 0.0f);
        if (A.dyn) { adjV[meta.x] = f4(bar.vA, 0.0f); adjW[meta.x] = f4(bar.wA, 0.0f); }
        if (B.dyn) { adjV[meta.y] = f4(bar.vB, 0.0f); adjW[meta.y] = f4(bar.wB, 0.0f); }
    };
    for (int it = sweeps - 1; it >= 0; --it)
        for (uint32_t r = runs; r-- > 0;) {
            for (uint32_t i = runStart[r] + threadIdx.x; i < runStart[r + 1]; i += kDiffThreads) visit(i, it, false);
            __syncthreads();
        }
    if (warm)
        for (uint32_t r = runs; r-- > 0;) {
            for (uint32_t i = runStart[r] + threadIdx.x; i < runStart[r + 1]; i += kDiffThreads) visit(i, sweeps, true);
            __syncthreads();
        }
}

__device__ inline void atomicAdd3(float4* p, vxadj::V<float> v) {
    atomicAdd(&p->x, v.x);
    atomicAdd(&p->y, v.y);
    atomicAdd(&p->z, v.z);
}
__device__ inline void atomicAddQ(float4* p, vxadj::Q<float> q) {
    atomicAdd(&p->x, q.x);
    atomicAdd(&p->y, q.y);
    atomicAdd(&p->z, q.z);
    atomicAdd(&p->w, q.w);
}

// Per-body parameter adjoints: (solver inverse mass, inverse inertia,
// friction, unused).
// The row prefix's adjoint of one substep, per contact: the accumulated
// row adjoints become position, orientation, inverse-mass, inverse-inertia
// and friction adjoints of the two bodies (atomic adds), and are cleared.
__global__ __launch_bounds__(kThreads) void kDiffPrefixAdj(SolveArrays sa, const vx_contact* __restrict__ cs, uint32_t m,
                                                           const vx_body* __restrict__ bodies, float h,
                                                           float4* __restrict__ rowBar, float4* __restrict__ adjX,
                                                           float4* __restrict__ adjQ, float4* __restrict__ adjParam) {
    uint32_t p = blockIdx.x * kThreads + threadIdx.x;
    if (p >= m) return;
    vxadj::Row<float> R;
    vxadj::Side<float> A, B;
    uint4 meta;
    diffRow(sa, p, h, R, A, B, meta);
    vxadj::RowBar<float> rb;
    loadBar(rowBar + 4 * (size_t)p, rb);
    for (int q = 0; q < 4; ++q) rowBar[4 * (size_t)p + q] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    if (meta.z & kSensorRow) return;
    const float mu1b = rb.mu1, mu2b = rb.mu2;
    vxadj::V<float> xAb{0.0f, 0.0f, 0.0f}, xBb{0.0f, 0.0f, 0.0f};
    vxadj::rowPrefixAdj(R, A, B, rb, xAb, xBb);
    const vx_contact& c = cs[p];
    const float fricBar = mu1b * R.mu1 + mu2b * R.mu2;  // d(mu_pair s)/d mu_body = (mu_pair s) / (2 mu_body)
    for (int side = 0; side < 2; ++side) {
        const bool dyn = side ? B.dyn : A.dyn;
        if (!dyn) continue;
        const uint32_t body = side ? meta.y : meta.x;
        const vx_body& b = bodies[body];
        atomicAdd3(adjX + body, side ? xBb : xAb);
        const vxadj::V<float> rbar = side ? rb.rb : rb.ra;
        if (b.shape != VX_SHAPE_SPHERE) {
            vxadj::Q<float> qb{0.0f, 0.0f, 0.0f, 0.0f};
            const vx_vec3 l = side ? c.localAnchorB : c.localAnchorA;
            vxadj::rotateAdjQ(vxadj::Q<float>{b.orientation.x, b.orientation.y, b.orientation.z, b.orientation.w},
                              vxadj::V<float>{l.x, l.y, l.z}, rbar, qb);
            atomicAddQ(adjQ + body, qb);
        }
        atomicAdd(&adjParam[body].x, side ? rb.imB : rb.imA);
        atomicAdd(&adjParam[body].y, side ? rb.iiB : rb.iiA);
        if (b.friction > 0.0f) atomicAdd(&adjParam[body].z, fricBar / (2.0f * b.friction));
    }
}

// Transform advance of one substep in reverse (isotropic branch), per body:
// `pre` holds the poses before the advance and the velocities it used.
__global__ __launch_bounds__(kThreads) void kDiffTransformAdj(const vx_body* __restrict__ pre, uint32_t n, float h,
                                                              float4* __restrict__ adjX, float4* __restrict__ adjQ,
                                                              float4* __restrict__ adjV, float4* __restrict__ adjW) {
    uint32_t i = blockIdx.x * kThreads + threadIdx.x;
    if (i >= n) return;
    const vx_body b = pre[i];
    if (isStatic(b) || isLocked(b) || asleepFully(b)) return;
    const float4 xb1 = adjX[i], qb1 = adjQ[i];
    vxadj::V<float> xb{0.0f, 0.0f, 0.0f}, vb = av(adjV[i]), wb = av(adjW[i]);
    vxadj::Q<float> qb{0.0f, 0.0f, 0.0f, 0.0f};
    vxadj::transformAdj(vxadj::Q<float>{b.orientation.x, b.orientation.y, b.orientation.z, b.orientation.w},
                        vxadj::V<float>{b.angularVelocity.x, b.angularVelocity.y, b.angularVelocity.z}, h, av(xb1),
                        vxadj::Q<float>{qb1.x, qb1.y, qb1.z, qb1.w}, xb, qb, vb, wb);
    adjX[i] = f4(xb, 0.0f);
    adjQ[i] = make_float4(qb.x, qb.y, qb.z, qb.w);
    adjV[i] = f4(vb, 0.0f);
    adjW[i] = f4(wb, 0.0f);
}

// The step-start renormalisation of the orientation in reverse.
__global__ __launch_bounds__(kThreads) void kDiffNormalizeAdj(const vx_body* __restrict__ start, uint32_t n,
                                                              float4* __restrict__ adjQ) {
    uint32_t i = blockIdx.x * kThreads + threadIdx.x;
    if (i >= n || isStatic(start[i])) return;
    const vx_quat q = start[i].orientation;
    float4 a = adjQ[i];
    vxadj::Q<float> qb{a.x, a.y, a.z, a.w};
    vxadj::normalizeAdj(vxadj::Q<float>{q.x, q.y, q.z, q.w}, qb);
    adjQ[i] = make_float4(qb.x, qb.y, qb.z, qb.w);
}

// Positional correction replayed per world (positionRow, world.cpp:4364-4397)
// with a tape of (g, fired) per visit.
__global__ __launch_bounds__(kDiffThreads) void kDiffCorrection(vx_body* __restrict__ bodies,
                                                                const vx_contact* __restrict__ cs,
                                                                const uint32_t* __restrict__ list,
                                                                const uint32_t* __restrict__ wStart,
                                                                const uint32_t* __restrict__ colorOf, uint32_t m,
                                                                int passes, float h, vx_solver_options opt,
                                                                float2* __restrict__ tape, uint32_t* __restrict__ overflow) {
    __shared__ uint32_t runStart[kDiffMaxRuns + 1];
    const uint32_t w = blockIdx.x, s0 = wStart[w], s1 = wStart[w + 1];
    if (s0 == s1) return;
    const uint32_t runs = diffRuns(list, colorOf, s0, s1, runStart, overflow);
    for (int it = 0; it < passes; ++it)
        for (uint32_t r = 0; r < runs; ++r) {
            for (uint32_t i = runStart[r] + threadIdx.x; i < runStart[r + 1]; i += kDiffThreads) {
                const uint32_t p = list[i];
                const vx_contact& c = cs[p];
                const vx_body a = bodies[c.a], b = bodies[c.b];
                const float g = liveGap(a, b, c);
                const bool fired = !(isSensor(a) || isSensor(b)) && solverInvMass(a) + solverInvMass(b) > 0.0f &&
                                   !(g >= -1e-3f);
                tape[(size_t)it * m + i] = make_float2(g, fired ? 1.0f : 0.0f);
                positionRow(bodies, cs, p, h, opt);
            }
            __syncthreads();
        }
}

__global__ __launch_bounds__(kDiffThreads) void kDiffCorrectionAdj(const vx_body* __restrict__ bodies,
                                                                   const vx_contact* __restrict__ cs,
                                                                   const uint32_t* __restrict__ list,
                                                                   const uint32_t* __restrict__ wStart,
                                                                   const uint32_t* __restrict__ colorOf, uint32_t m,
                                                                   int passes, const float2* __restrict__ tape,
                                                                   float4* __restrict__ adjX, float4* __restrict__ adjQ,
                                                                   float4* __restrict__ adjParam,
                                                                   uint32_t* __restrict__ overflow) {
    __shared__ uint32_t runStart[kDiffMaxRuns + 1];
    const uint32_t w = blockIdx.x, s0 = wStart[w], s1 = wStart[w + 1];
    if (s0 == s1) return;
    const uint32_t runs = diffRuns(list, colorOf, s0, s1, runStart, overflow);
    for (int it = passes - 1; it >= 0; --it)
        for (uint32_t r = runs; r-- > 0;) {
            for (uint32_t i = runStart[r] + threadIdx.x; i < runStart[r + 1]; i += kDiffThreads) {
                const float2 t = tape[(size_t)it * m + i];
                if (t.y == 0.0f) continue;
                const vx_contact& c = cs[list[i]];
                const vx_body& a = bodies[c.a];
                const vx_body& b = bodies[c.b];
                const bool dynA = isDynamic(a), dynB = isDynamic(b);
                const vxadj::V<float> n{c.normal.x, c.normal.y, c.normal.z};
                float imAb = 0.0f, imBb = 0.0f;
                const float gb = vxadj::correctionAdj(t.x, solverInvMass(a), solverInvMass(b), dynA, dynB, n,
                                                      av(adjX[c.a]), av(adjX[c.b]), imAb, imBb);
                for (int side = 0; side < 2; ++side) {
                    if (!(side ? dynB : dynA)) continue;
                    const uint32_t body = side ? c.b : c.a;
                    const vx_body& bd = side ? b : a;
                    const float sgn = side ? -1.0f : 1.0f;
                    float4 x = adjX[body];
                    adjX[body] = make_float4(x.x + sgn * n.x * gb, x.y + sgn * n.y * gb, x.z + sgn * n.z * gb, 0.0f);
                    if (bd.shape != VX_SHAPE_SPHERE) {
                        vxadj::Q<float> qb{0.0f, 0.0f, 0.0f, 0.0f};
                        const vx_vec3 l = side ? c.localAnchorB : c.localAnchorA;
                        vxadj::rotateAdjQ(
                            vxadj::Q<float>{bd.orientation.x, bd.orientation.y, bd.orientation.z, bd.orientation.w},
                            vxadj::V<float>{l.x, l.y, l.z}, n * (sgn * gb), qb);
                        float4 q = adjQ[body];
                        adjQ[body] = make_float4(q.x + qb.x, q.y + qb.y, q.z + qb.z, q.w + qb.w);
                    }
                    float4 pp = adjParam[body];
                    pp.x += side ? imBb : imAb;
                    adjParam[body] = pp;
                }
            }
            __syncthreads();
        }
}

// lamBarPrev[match[k]] += lamBar[k]: the warm-start impulses of this step are
// the matched contacts' final impulses of the previous step.
__global__ __launch_bounds__(kThreads) void kDiffLambdaCarry(const uint32_t* __restrict__ match, uint32_t m,
                                                             const float4* __restrict__ lamBar,
                                                             float4* __restrict__ lamBarPrev) {
    uint32_t k = blockIdx.x * kThreads + threadIdx.x;
    if (k >= m) return;
    const uint32_t j = match[k];
    if (j == 0xffffffffu) return;
    const float4 b = lamBar[k];
    atomicAdd(&lamBarPrev[j].x, b.x);
    atomicAdd(&lamBarPrev[j].y, b.y);
    atomicAdd(&lamBarPrev[j].z, b.z);
}

// Per world: centre of mass of the dynamic bodies [b0, b1) and the loss
// |com - target|^2; with lossBar, seeds the position and mass adjoints.
__global__ __launch_bounds__(kThreads) void kDiffLoss(const vx_body* __restrict__ bodies,
                                                      const uint32_t* __restrict__ bodyStart, uint32_t W,
                                                      const float* __restrict__ target, float* __restrict__ loss,
                                                      float4* __restrict__ com, const float* __restrict__ lossBar,
                                                      float4* __restrict__ adjX, float* __restrict__ adjMass) {
    uint32_t w = blockIdx.x * kThreads + threadIdx.x;
    if (w >= W) return;
    double M = 0.0, cx = 0.0, cy = 0.0, cz = 0.0;
    for (uint32_t i = bodyStart[w]; i < bodyStart[w + 1]; ++i) {
        const vx_body& b = bodies[i];
        if (!isDynamic(b) || b.invMass <= 0.0f) continue;
        const double mass = 1.0 / (double)b.invMass;
        M += mass;
        cx += mass * b.position.x; cy += mass * b.position.y; cz += mass * b.position.z;
    }
    if (M <= 0.0) M = 1.0;
    const float c[3] = {(float)(cx / M), (float)(cy / M), (float)(cz / M)};
    const float d[3] = {c[0] - target[3 * w], c[1] - target[3 * w + 1], c[2] - target[3 * w + 2]};
    if (loss) loss[w] = d[0] * d[0] + d[1] * d[1] + d[2] * d[2];
    if (com) com[w] = make_float4(c[0], c[1], c[2], 0.0f);
    if (!lossBar) return;
    const float lb = lossBar[w];
    for (uint32_t i = bodyStart[w]; i < bodyStart[w + 1]; ++i) {
        const vx_body& b = bodies[i];
        if (!isDynamic(b) || b.invMass <= 0.0f) continue;
        const float mass = 1.0f / b.invMass, k = 2.0f * lb * mass / (float)M;
        adjX[i] = make_float4(adjX[i].x + k * d[0], adjX[i].y + k * d[1], adjX[i].z + k * d[2], 0.0f);
        adjMass[i] += 2.0f * lb / (float)M *
                      (d[0] * (b.position.x - c[0]) + d[1] * (b.position.y - c[1]) + d[2] * (b.position.z - c[2]));
    }
}

// Parameter gradients from the adjoints at the start of the rollout:
// velocities, friction, and mass (through 1/m and the inverse inertia k/m).
__global__ __launch_bounds__(kThreads) void kDiffParams(const vx_body* __restrict__ bodies, uint32_t n,
                                                        const float4* __restrict__ adjV, const float4* __restrict__ adjW,
                                                        const float4* __restrict__ adjParam,
                                                        const float* __restrict__ adjMass, float* __restrict__ gV,
                                                        float* __restrict__ gW, float* __restrict__ gFriction,
                                                        float* __restrict__ gMass) {
    uint32_t i = blockIdx.x * kThreads + threadIdx.x;
    if (i >= n) return;
    const vx_body& b = bodies[i];
    const bool dyn = isDynamic(b) && b.invMass > 0.0f;
    const float4 v = adjV[i], w = adjW[i], pp = adjParam[i];
    gV[3 * i] = dyn ? v.x : 0.0f; gV[3 * i + 1] = dyn ? v.y : 0.0f; gV[3 * i + 2] = dyn ? v.z : 0.0f;
    gW[3 * i] = dyn ? w.x : 0.0f; gW[3 * i + 1] = dyn ? w.y : 0.0f; gW[3 * i + 2] = dyn ? w.z : 0.0f;
    gFriction[i] = dyn ? pp.z : 0.0f;
    const float im = b.invMass, ii = b.invInertia.x;
    gMass[i] = dyn ? adjMass[i] - pp.x * im * im - pp.y * ii * im : 0.0f;
}

}  // namespace vxk
