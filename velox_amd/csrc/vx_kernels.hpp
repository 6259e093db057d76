// Device kernels of the HIP device step (one body, pair or contact per lane).
#pragma once
#include "vxd_solver.hpp"

namespace vxk {

using namespace vxd;

constexpr int kThreads = 256;
constexpr uint32_t kNone = 0xffffffffu;
constexpr int kMaxPairContacts = kMaxContactsPerPair;  // 12 (narrowphase.h:12)

struct Grid {           // dense cell grid of the broad phase
    float cell;         // cell edge
    int minx, miny, minz;
    uint32_t nx, ny, nz;
    uint64_t ncells;    // cells of one world's grid
    uint64_t sentinel;  // key of oversize bodies: worlds * ncells
};

__device__ inline float4 f4(V3 v, float w = 0.0f) { return make_float4(v.x, v.y, v.z, w); }
__device__ inline V3 v3f(float4 v) { return {v.x, v.y, v.z}; }

// ------------------------------------------------------------- bodies --
// world.cpp:3975 (renormalise), solver.cpp:258-270 (first-substep velocity),
// world.cpp:4186-4188 (pre-step pose for the CCD net).
__global__ __launch_bounds__(kThreads) void kBeginStep(vx_body* __restrict__ bodies, uint32_t n, V3 g, float h,
                                                       vx_vec3* __restrict__ prevPos, vx_quat* __restrict__ prevQ,
                                                       int integrate) {
    uint32_t i = blockIdx.x * kThreads + threadIdx.x;
    if (i >= n) return;
    vx_body b = bodies[i];
    b.orientation = raw(qnormalize(q4(b.orientation)));
    if (integrate) integrateBody(b, g, h);
    bodies[i].orientation = b.orientation;
    bodies[i].velocity = b.velocity;
    bodies[i].angularVelocity = b.angularVelocity;
    if (prevPos) { prevPos[i] = b.position; prevQ[i] = b.orientation; }
}

__global__ __launch_bounds__(kThreads) void kIntegrate(vx_body* __restrict__ bodies, uint32_t n, V3 g, float h) {
    uint32_t i = blockIdx.x * kThreads + threadIdx.x;
    if (i >= n) return;
    vx_body b = bodies[i];
    if (!isDynamic(b) || isLocked(b) || asleepFully(b)) return;
    integrateBody(b, g, h);
    bodies[i].velocity = b.velocity;
    bodies[i].angularVelocity = b.angularVelocity;
}

// world.cpp:4219-4222 / backend_cuda.cu:91-97
__global__ __launch_bounds__(kThreads) void kTransforms(vx_body* __restrict__ bodies, uint32_t n, float h) {
    uint32_t i = blockIdx.x * kThreads + threadIdx.x;
    if (i >= n) return;
    vx_body b = bodies[i];
    if (isStatic(b) || isLocked(b) || asleepFully(b)) return;
    advanceTransform(b, h);
    bodies[i].position = b.position;
    bodies[i].orientation = b.orientation;
    bodies[i].angularVelocity = b.angularVelocity;
}

// world.cpp:4493-4496
__global__ __launch_bounds__(kThreads) void kClearForces(vx_body* __restrict__ bodies, uint32_t n) {
    uint32_t i = blockIdx.x * kThreads + threadIdx.x;
    if (i >= n) return;
    bodies[i].force = {0, 0, 0};
    bodies[i].torque = {0, 0, 0};
}

// ------------------------------------------------------------ broad phase --
__device__ inline bool oversized(const vx_body& b, float extent, float cell) {
    return extent > cell || b.shape == VX_SHAPE_PLANE || b.shape == VX_SHAPE_MESH;
}

// narrowphase.h:369-398; planes unbounded, meshes use their root box
// (backend_cuda.cu:103-129).  The largest finite extent sets the cell edge.
// Extent histogram of the grid-cell choice: 1/8-octave buckets of the float
// bits, 2^-16 to 2^16 (the ends clamped).
constexpr uint32_t kExtBuckets = 256, kExtBucketBase = (127u - 16u) << 3;
__device__ inline uint32_t extBucket(float e) {
    const int b = (int)(__float_as_uint(e) >> 20) - (int)kExtBucketBase;
    return (uint32_t)(b < 0 ? 0 : b > (int)kExtBuckets - 1 ? (int)kExtBuckets - 1 : b);
}
__device__ inline float extBucketEnd(uint32_t b) { return __uint_as_float((b + kExtBucketBase + 1) << 20); }
static_assert(kThreads == (int)kExtBuckets, "kAabb keeps one histogram bucket per thread");

__global__ __launch_bounds__(kThreads) void kAabb(const vx_body* __restrict__ bodies, uint32_t n, Soup soup, float dt,
                                                  float4* __restrict__ lo, float4* __restrict__ hi,
                                                  uint32_t* __restrict__ extentBits, uint32_t* __restrict__ extHist) {
    __shared__ uint32_t hist[kExtBuckets];
    hist[threadIdx.x] = 0;
    __syncthreads();
    uint32_t i = blockIdx.x * kThreads + threadIdx.x;
    float ext = 0.0f;
    if (i < n) {
        const vx_body b = bodies[i];
        V3 l, h;
        if (b.shape == VX_SHAPE_PLANE) { l = {-1e30f, -1e30f, -1e30f}; h = {1e30f, 1e30f, 1e30f}; }
        else if (b.shape == VX_SHAPE_MESH) {
            const vx_mesh& m = soup.meshes[b.meshIndex];
            l = v3(m.aabbMin); h = v3(m.aabbMax);
        } else {
            bodyAabb(b, dt, l, h);
            float e = fmaxf(h.x - l.x, fmaxf(h.y - l.y, h.z - l.z));
            if (e < 1e6f) ext = e;
        }
        // lo.w marks the shapes that are always oversize (planes, meshes).
        lo[i] = f4(l, b.shape == VX_SHAPE_PLANE || b.shape == VX_SHAPE_MESH ? 1.0f : 0.0f);
        hi[i] = f4(h);
    }
    if (ext > 0.0f) atomicAdd(&hist[extBucket(ext)], 1u);
    // One atomic per wavefront: positive floats order like their bit patterns.
    for (int off = 32; off > 0; off >>= 1) ext = fmaxf(ext, __shfl_xor(ext, off));
    if ((threadIdx.x & 63) == 0 && ext > 0.0f) atomicMax(extentBits, __float_as_uint(ext));
    __syncthreads();
    if (hist[threadIdx.x]) atomicAdd(&extHist[threadIdx.x], hist[threadIdx.x]);
}

// The grid cell's extent: the largest body extent, unless a few bodies
// (spinning ones, whose swept boxes grow with their angular speed) are far
// larger than the rest.  Then the cell is the end of the smallest extent
// bucket with at most 32 + n/2048 bodies above it, and those bodies take the
// oversize path (each tested against every body): the pair set is the same
// whatever the cell, since every body larger than the cell is oversize.
// Clears the histogram for the next step.  One workgroup of kExtBuckets.
__global__ __launch_bounds__(kExtBuckets) void kCellSize(uint32_t* __restrict__ extHist, uint32_t n,
                                                         uint32_t* __restrict__ extentBits) {
    __shared__ uint32_t above[kExtBuckets + 1];  // above[b]: bodies in buckets b and later
    __shared__ uint32_t pick;
    const uint32_t t = threadIdx.x;
    above[t] = extHist[t];
    extHist[t] = 0;
    if (t == 0) { above[kExtBuckets] = 0; pick = kExtBuckets; }
    __syncthreads();
    for (uint32_t off = 1; off < kExtBuckets; off <<= 1) {
        const uint32_t v = t + off < kExtBuckets ? above[t + off] : 0u;
        __syncthreads();
        above[t] += v;
        __syncthreads();
    }
    // The smallest bucket with at most `limit` bodies in later buckets.
    const uint32_t limit = 32u + n / 2048u;
    const bool ok = above[t + 1] <= limit, okPrev = t > 0 && above[t] <= limit;
    if (ok && !okPrev) pick = t;
    __syncthreads();
    if (t != 0 || pick >= kExtBuckets) return;
    const float maxExt = __uint_as_float(*extentBits);
    const float edge = extBucketEnd(pick);
    if (maxExt > 1.5f * edge) *extentBits = __float_as_uint(edge);
}

__device__ inline float gridCellOf(const uint32_t* extentBits) {
    float c = __uint_as_float(*extentBits);
    // The grid edge sits just above the largest extent so that overlapping
    // boxes always land in neighbouring cells despite rounding.
    return c > 0.0f ? c * 1.0001f : 1.0f;
}

__device__ inline void cellOf(float4 lo, float4 hi, float inv, int& cx, int& cy, int& cz) {
    cx = (int)floorf((lo.x + hi.x) * 0.5f * inv);
    cy = (int)floorf((lo.y + hi.y) * 0.5f * inv);
    cz = (int)floorf((lo.z + hi.z) * 0.5f * inv);
}

// Reduced over the wavefront first: one set of atomics per wavefront (the
// bounds are six words every body would otherwise hit).
__global__ __launch_bounds__(kThreads) void kCellBounds(const float4* __restrict__ lo,
                                                        const float4* __restrict__ hi, uint32_t n,
                                                        const uint32_t* __restrict__ extentBits, int* __restrict__ bounds,
                                                        uint32_t* __restrict__ oversizeCount) {
    const uint32_t i = blockIdx.x * kThreads + threadIdx.x;
    bool over = false, in = false;
    int cx = 0x7fffffff, cy = 0x7fffffff, cz = 0x7fffffff, mx = (int)0x80000000, my = (int)0x80000000,
        mz = (int)0x80000000;
    if (i < n) {
        const float cell = gridCellOf(extentBits);
        float4 l = lo[i], h = hi[i];
        float ext = fmaxf(h.x - l.x, fmaxf(h.y - l.y, h.z - l.z));
        over = ext > cell || l.w != 0.0f;
        if (!over) {
            cellOf(l, h, 1.0f / cell, cx, cy, cz);
            mx = cx; my = cy; mz = cz;
            in = true;
        }
    }
    for (int off = 32; off > 0; off >>= 1) {
        cx = min(cx, __shfl_xor(cx, off)); cy = min(cy, __shfl_xor(cy, off)); cz = min(cz, __shfl_xor(cz, off));
        mx = max(mx, __shfl_xor(mx, off)); my = max(my, __shfl_xor(my, off)); mz = max(mz, __shfl_xor(mz, off));
    }
    const uint64_t overMask = __ballot(over), inMask = __ballot(in);
    if ((threadIdx.x & 63) == 0) {
        if (overMask) atomicAdd(oversizeCount, (uint32_t)__popcll(overMask));
        if (inMask) {
            atomicMin(&bounds[0], cx); atomicMin(&bounds[1], cy); atomicMin(&bounds[2], cz);
            atomicMax(&bounds[3], mx); atomicMax(&bounds[4], my); atomicMax(&bounds[5], mz);
        }
    }
}

// Batched worlds share one grid: a body's key is world * ncells + cell, so
// neighbourhood searches never leave the body's own world.
__global__ __launch_bounds__(kThreads) void kCellKeys(const vx_body* __restrict__ bodies, const float4* __restrict__ lo,
                                                      const float4* __restrict__ hi, uint32_t n, Grid g,
                                                      const uint32_t* __restrict__ worldOf,
                                                      uint64_t* __restrict__ keys, uint32_t* __restrict__ idx,
                                                      uint32_t* __restrict__ oversize, uint32_t* __restrict__ oversizeCount,
                                                      uint32_t* __restrict__ worldOversize) {
    uint32_t i = blockIdx.x * kThreads + threadIdx.x;
    if (i >= n) return;
    idx[i] = i;
    float4 l = lo[i], h = hi[i];
    float ext = fmaxf(h.x - l.x, fmaxf(h.y - l.y, h.z - l.z));
    const uint8_t shape = bodies[i].shape;
    if (ext > g.cell || shape == VX_SHAPE_PLANE || shape == VX_SHAPE_MESH) {
        keys[i] = g.sentinel;
        uint32_t slot = atomicAdd(oversizeCount, 1u);
        oversize[slot] = i;
        if (worldOf) atomicAdd(&worldOversize[worldOf[i]], 1u);
        return;
    }
    int cx, cy, cz;
    cellOf(l, h, 1.0f / g.cell, cx, cy, cz);
    const uint64_t world = worldOf ? worldOf[i] : 0u;
    keys[i] = world * g.ncells +
              ((uint64_t)(cx - g.minx) * g.ny + (uint64_t)(cy - g.miny)) * g.nz + (uint64_t)(cz - g.minz);
}

__device__ inline bool pairAllowed(const vx_body& a, const vx_body& b) {
    if (inert(a) && inert(b)) return false;
    return canCollide(a, b);
}

// Contact key of a pair: plane/mesh pairs name the convex body first
// (narrowphase.h:665-671), packed as (first << nb) | second.
__device__ inline uint64_t packPair(const vx_body* bodies, uint32_t i, uint32_t j, int nb) {
    uint32_t lo = i < j ? i : j, hi = i < j ? j : i;
    uint8_t s = bodies[lo].shape;
    if (s == VX_SHAPE_PLANE || s == VX_SHAPE_MESH) return ((uint64_t)hi << nb) | lo;
    return ((uint64_t)lo << nb) | hi;
}

__device__ inline uint32_t lowerBound(const uint64_t* keys, uint32_t n, uint64_t key) {
    uint32_t lo = 0, hi = n;
    while (lo < hi) {
        uint32_t mid = (lo + hi) >> 1;
        if (keys[mid] < key) lo = mid + 1; else hi = mid;
    }
    return lo;
}

// Inclusive wavefront prefix sum of a per-lane count.
__device__ inline uint32_t waveInclusive(uint32_t v) {
    const int lane = threadIdx.x & 63;
    for (int off = 1; off < 64; off <<= 1) {
        uint32_t t = __shfl_up(v, off);
        if (lane >= off) v += t;
    }
    return v;
}

// Reserves `count` slots for this lane with one atomic per wavefront.
__device__ inline uint32_t waveReserve(uint32_t count, uint32_t* counter) {
    uint32_t incl = waveInclusive(count);
    uint32_t total = __shfl(incl, 63);
    uint32_t base = 0;
    if ((threadIdx.x & 63) == 63 && total) base = atomicAdd(counter, total);
    base = __shfl(base, 63);
    return base + incl - count;
}

// 27-neighbourhood candidates of each non-oversize body in sorted order,
// exact AABB test, inert/filter rules of backend_cuda.cu:174-218; each pair
// found once (from its lower body index).  Two passes: count, then write.
template <bool kWrite>
__device__ inline uint32_t gridVisit(const vx_body* __restrict__ bodies, const float4* __restrict__ lo,
                                     const float4* __restrict__ hi, const uint64_t* __restrict__ keys,
                                     const uint32_t* __restrict__ idx, const uint2* __restrict__ cellRange, uint32_t n,
                                     const Grid& g, uint32_t s, int nb, uint64_t* out, uint32_t at, uint32_t cap) {
    const uint64_t wkey = keys[s];
    const uint64_t base = (wkey / g.ncells) * g.ncells;  // the body's world
    const uint64_t key = wkey - base;
    const uint32_t i = idx[s];
    const uint32_t z = (uint32_t)(key % g.nz);
    const uint32_t y = (uint32_t)((key / g.nz) % g.ny);
    const uint32_t x = (uint32_t)(key / ((uint64_t)g.nz * g.ny));
    const float4 li = lo[i], hiI = hi[i];
    const vx_body& a = bodies[i];
    uint32_t count = 0;
    for (int dx = -1; dx <= 1; ++dx) {
        if ((dx < 0 && x == 0) || (uint32_t)(x + dx) >= g.nx) continue;
        for (int dy = -1; dy <= 1; ++dy) {
            if ((dy < 0 && y == 0) || (uint32_t)(y + dy) >= g.ny) continue;
            for (int dz = -1; dz <= 1; ++dz) {
                if ((dz < 0 && z == 0) || (uint32_t)(z + dz) >= g.nz) continue;
                const uint64_t nk = base + ((uint64_t)(x + dx) * g.ny + (y + dy)) * g.nz + (z + dz);
                uint32_t t0, t1;
                if (cellRange) {
                    const uint2 r = cellRange[nk];
                    t0 = r.x;
                    t1 = r.y;
                } else {
                    t0 = lowerBound(keys, n, nk);
                    t1 = t0;
                    while (t1 < n && keys[t1] == nk) ++t1;
                }
                for (uint32_t t = t0; t < t1; ++t) {
                    const uint32_t j = idx[t];
                    if (j <= i) continue;
                    const float4 lj = lo[j], hj = hi[j];
                    if (!(hiI.x >= lj.x && li.x <= hj.x && hiI.y >= lj.y && li.y <= hj.y && hiI.z >= lj.z && li.z <= hj.z))
                        continue;
                    if (!pairAllowed(a, bodies[j])) continue;
                    if (kWrite && at + count < cap) out[at + count] = ((uint64_t)i << nb) | j;
                    ++count;
                }
            }
        }
    }
    return count;
}

__global__ __launch_bounds__(kThreads) void kGridPairs(const vx_body* __restrict__ bodies, const float4* __restrict__ lo,
                                                       const float4* __restrict__ hi, const uint64_t* __restrict__ keys,
                                                       const uint32_t* __restrict__ idx,
                                                       const uint2* __restrict__ cellRange, uint32_t n, Grid g, int nb,
                                                       uint64_t* __restrict__ pairs, uint32_t* __restrict__ pairCount,
                                                       uint32_t cap) {
    uint32_t s = blockIdx.x * kThreads + threadIdx.x;
    const bool valid = s < n && keys[s] != g.sentinel;
    uint32_t cnt = valid ? gridVisit<false>(bodies, lo, hi, keys, idx, cellRange, n, g, s, nb, nullptr, 0, 0) : 0u;
    uint32_t at = waveReserve(cnt, pairCount);
    if (valid && cnt) gridVisit<true>(bodies, lo, hi, keys, idx, cellRange, n, g, s, nb, pairs, at, cap);
}

// Sorted-key run of every occupied cell ([start, end) into the sorted
// arrays); empty cells keep (0, 0).  Replaces the binary searches of the
// neighbourhood walk when the cell count is small enough for a table.
__global__ __launch_bounds__(kThreads) void kCellRanges(const uint64_t* __restrict__ keys, uint32_t n, uint64_t sentinel,
                                                        uint2* __restrict__ cellRange) {
    uint32_t t = blockIdx.x * kThreads + threadIdx.x;
    if (t >= n) return;
    const uint64_t key = keys[t];
    if (key == sentinel) return;
    if (t == 0 || keys[t - 1] != key) cellRange[key].x = t;
    if (t + 1 == n || keys[t + 1] != key) cellRange[key].y = t + 1;
}

// Oversize bodies (planes, meshes, huge sweeps) against every body
// (backend_cuda.cu:221-250).
__global__ __launch_bounds__(kThreads) void kOversizePairs(const vx_body* __restrict__ bodies, const float4* __restrict__ lo,
                                                           const float4* __restrict__ hi,
                                                           const uint32_t* __restrict__ oversize, uint32_t oversizeCount,
                                                           uint32_t n, Grid g, int nb, uint64_t* __restrict__ pairs,
                                                           uint32_t* __restrict__ pairCount, uint32_t cap) {
    uint64_t flat = (uint64_t)blockIdx.x * kThreads + threadIdx.x;
    bool hit = false;
    uint64_t key = 0;
    if (flat < (uint64_t)oversizeCount * n) {
        uint32_t o = (uint32_t)(flat / n), j = (uint32_t)(flat % n);
        uint32_t i = oversize[o];
        if (j != i) {
            float4 lj = lo[j], hj = hi[j];
            float extJ = fmaxf(hj.x - lj.x, fmaxf(hj.y - lj.y, hj.z - lj.z));
            bool jOver = oversized(bodies[j], extJ, g.cell);
            float4 li = lo[i], hiI = hi[i];
            if (!(jOver && j < i) &&
                hiI.x >= lj.x && li.x <= hj.x && hiI.y >= lj.y && li.y <= hj.y && hiI.z >= lj.z && li.z <= hj.z &&
                pairAllowed(bodies[i], bodies[j])) {
                hit = true;
                key = packPair(bodies, i, j, nb);
            }
        }
    }
    uint32_t at = waveReserve(hit ? 1u : 0u, pairCount);
    if (hit && at < cap) pairs[at] = key;
}

// Batched worlds: oversize bodies listed per world, then each body tested
// against its own world's oversize bodies only (the same tests and the same
// duplicate rule as kOversizePairs).
__global__ __launch_bounds__(kThreads) void kOversizeByWorld(const uint32_t* __restrict__ oversize, uint32_t oversizeCount,
                                                             const uint32_t* __restrict__ worldOf,
                                                             const uint32_t* __restrict__ worldStart,
                                                             uint32_t* __restrict__ worldFill,
                                                             uint32_t* __restrict__ byWorld) {
    uint32_t o = blockIdx.x * kThreads + threadIdx.x;
    if (o >= oversizeCount) return;
    const uint32_t i = oversize[o], w = worldOf[i];
    byWorld[worldStart[w] + atomicAdd(&worldFill[w], 1u)] = i;
}

__global__ __launch_bounds__(kThreads) void kOversizePairsWorld(const vx_body* __restrict__ bodies,
                                                                const float4* __restrict__ lo,
                                                                const float4* __restrict__ hi,
                                                                const uint32_t* __restrict__ worldOf,
                                                                const uint32_t* __restrict__ worldStart,
                                                                const uint32_t* __restrict__ byWorld, uint32_t n,
                                                                Grid g, int nb, uint64_t* __restrict__ pairs,
                                                                uint32_t* __restrict__ pairCount, uint32_t cap) {
    const uint32_t j = blockIdx.x * kThreads + threadIdx.x;
    uint32_t first = 0, last = 0;
    float4 lj{}, hj{};
    bool jOver = false;
    if (j < n) {
        const uint32_t w = worldOf[j];
        first = worldStart[w];
        last = worldStart[w + 1];
        lj = lo[j];
        hj = hi[j];
        const float extJ = fmaxf(hj.x - lj.x, fmaxf(hj.y - lj.y, hj.z - lj.z));
        jOver = oversized(bodies[j], extJ, g.cell);
    }
    // Lanes of a wavefront walk their own lists; emission is reserved per hit.
    for (uint32_t o = first; o < last; ++o) {
        const uint32_t i = byWorld[o];
        const float4 li = lo[i], hiI = hi[i];
        const bool hit = j != i && !(jOver && j < i) &&
                         hiI.x >= lj.x && li.x <= hj.x && hiI.y >= lj.y && li.y <= hj.y && hiI.z >= lj.z &&
                         li.z <= hj.z && pairAllowed(bodies[i], bodies[j]);
        if (hit) {
            const uint32_t at = atomicAdd(pairCount, 1u);
            if (at < cap) pairs[at] = packPair(bodies, i, j, nb);
        }
    }
}

// ----------------------------------------------------------- narrow phase --
__device__ inline bool canonicalLess(const vx_contact& x, const vx_contact& y) {  // backend_cuda.cu:587-595
    if (x.point.x != y.point.x) return x.point.x < y.point.x;
    if (x.point.y != y.point.y) return x.point.y < y.point.y;
    return x.point.z < y.point.z;
}

// One pair per lane: collidePair, then the pair's contacts in point order
// (stable, so equal points keep emission order) into the pair's 12 slots.
// Shape class of every candidate pair (the lower body's shape, then the
// other's) and the identity order.  Sorted by class, this order puts pairs of
// the same shapes side by side in a wavefront of kNarrow and kCcdCheck, which
// then run one shape routine instead of all of them; every result is still
// written at the pair's own index.
__global__ __launch_bounds__(kThreads) void kPairClass(const vx_body* __restrict__ bodies,
                                                       const uint64_t* __restrict__ pairs, uint32_t np, int nb,
                                                       uint64_t* __restrict__ keys, uint32_t* __restrict__ vals) {
    uint32_t p = blockIdx.x * kThreads + threadIdx.x;
    if (p >= np) return;
    const uint64_t key = pairs[p];
    const uint32_t ca = (uint32_t)(key >> nb), cb = (uint32_t)(key & ((1ull << nb) - 1));
    const uint32_t i = ca < cb ? ca : cb, j = ca < cb ? cb : ca;
    // Classes in descending shape order (compounds, hulls, meshes, capsules,
    // boxes, planes, spheres), so the longest wavefronts are dispatched first
    // and the ones that wait for a free SIMD are the short ones.  Each pair's results land at its
    // own index, so the order is a schedule only.
    keys[p] = 255u - ((bodies[i].shape & 15u) << 4 | (bodies[j].shape & 15u));
    vals[p] = p;
}
constexpr int kPairClassBits = 8;

// The collision kernels run one wavefront per workgroup: at one wavefront
// per SIMD, a workgroup of several would hold its CU's slots until its
// slowest wavefront ends, and the second round of wavefronts could not start
// on a SIMD as soon as it frees.
constexpr int kCollideThreads = 64;

__global__ __launch_bounds__(kCollideThreads) void kNarrow(const vx_body* __restrict__ bodies, Soup soup, float dt,
                                                    const uint64_t* __restrict__ pairs, uint32_t np, int nb,
                                                    const uint32_t* __restrict__ order,
                                                    vx_contact* __restrict__ slots, uint32_t* __restrict__ counts) {
    const uint32_t q = blockIdx.x * kCollideThreads + threadIdx.x;
    if (q >= np) return;
    const uint32_t p = order[q];
    uint64_t key = pairs[p];
    uint32_t ca = (uint32_t)(key >> nb), cb = (uint32_t)(key & ((1ull << nb) - 1));
    uint32_t i = ca < cb ? ca : cb, j = ca < cb ? cb : ca;
    const vx_body A = bodies[i], B = bodies[j];
    vx_contact buf[kMaxPairContacts];
    int cnt = collidePair(A, B, i, j, soup, dt, buf, kMaxPairContacts);
    for (int k = 1; k < cnt; ++k) {
        vx_contact c = buf[k];
        int m = k - 1;
        while (m >= 0 && canonicalLess(c, buf[m])) { buf[m + 1] = buf[m]; --m; }
        buf[m + 1] = c;
    }
    vx_contact* dst = slots + (size_t)p * kMaxPairContacts;
    for (int k = 0; k < cnt; ++k) dst[k] = buf[k];
    counts[p] = (uint32_t)cnt;
}

__global__ __launch_bounds__(kThreads) void kCompact(const vx_contact* __restrict__ slots, const uint32_t* __restrict__ counts,
                                                     const uint32_t* __restrict__ offsets, uint32_t np,
                                                     vx_contact* __restrict__ out) {
    uint32_t p = blockIdx.x * kThreads + threadIdx.x;
    if (p >= np) return;
    uint32_t c = counts[p], at = offsets[p];
    for (uint32_t k = 0; k < c; ++k) out[at + k] = slots[(size_t)p * kMaxPairContacts + k];
}

// -------------------------------------------------------- warm starting --
// world.cpp:4109-4128.  Order-dependent (a body woken by one contact can wake
// the next), so one lane applies the rule in contact order; but only to
// contacts with a fully asleep body, which the workgroup finds 1024 at a
// time.  Bodies only ever wake here, so a stale "asleep" read in the parallel
// filter can only add a contact that the exact rule then leaves alone.
constexpr int kWakeThreads = 1024;
__global__ __launch_bounds__(kWakeThreads) void kWakePass(vx_body* bodies, const vx_contact* cs, uint32_t m) {
    __shared__ uint8_t relevant[kWakeThreads];
    for (uint32_t base = 0; base < m; base += kWakeThreads) {
        const uint32_t k = base + threadIdx.x;
        bool r = false;
        if (k < m) r = asleepFully(bodies[cs[k].a]) || asleepFully(bodies[cs[k].b]);
        relevant[threadIdx.x] = r ? 1 : 0;
        if (__syncthreads_or(r) && threadIdx.x == 0) {
            const uint32_t end = m - base < (uint32_t)kWakeThreads ? m - base : (uint32_t)kWakeThreads;
            for (uint32_t j = 0; j < end; ++j) {
                if (!relevant[j]) continue;
                const vx_contact& c = cs[base + j];
                vx_body& a = bodies[c.a];
                vx_body& b = bodies[c.b];
                if (isSensor(a) || isSensor(b)) continue;
                const bool impact = c.vn0 < -0.1f;
                const bool movingA = lengthSq(v3(a.velocity)) + lengthSq(v3(a.angularVelocity)) > 1e-3f;
                const bool movingB = lengthSq(v3(b.velocity)) + lengthSq(v3(b.angularVelocity)) > 1e-3f;
                if (asleepFully(a) && !isStatic(b) && !asleepFully(b) && (impact || movingB)) { a.asleep = 0; a.sleepTimer = 0.0f; }
                if (asleepFully(b) && !isStatic(a) && !asleepFully(a) && (impact || movingA)) { b.asleep = 0; b.sleepTimer = 0.0f; }
            }
        }
        __syncthreads();
    }
}

__device__ inline uint64_t pairKeyOf(const vx_contact& c) { return (uint64_t)c.a << 32 | c.b; }

__device__ inline void copyImpulses(vx_contact& c, const vx_contact& p) {
    c.normalImpulse = p.normalImpulse; c.tangentImpulse1 = p.tangentImpulse1; c.tangentImpulse2 = p.tangentImpulse2;
    c.rollingImpulse1 = p.rollingImpulse1; c.rollingImpulse2 = p.rollingImpulse2; c.spinningImpulse = p.spinningImpulse;
}

// world.cpp:4133-4184: feature-key match, vn0 carry, nearest point within 5 cm.
// With `prevPos` (each cache entry's position in the previous step's solve
// order), `match` receives the position of the entry whose impulses the
// contact took, or ~0 (the differentiable form's record of the warm start).
__global__ __launch_bounds__(kThreads) void kWarmMatch(vx_contact* __restrict__ cur, uint32_t m,
                                                       const vx_contact* __restrict__ prev, uint32_t mp,
                                                       const uint32_t* __restrict__ prevPos = nullptr,
                                                       uint32_t* __restrict__ match = nullptr) {
    uint32_t k = blockIdx.x * kThreads + threadIdx.x;
    if (k >= m || mp == 0) return;
    vx_contact c = cur[k];
    const uint64_t key = pairKeyOf(c);
    uint32_t lo = 0, hi = mp;
    while (lo < hi) { uint32_t mid = (lo + hi) >> 1; if (pairKeyOf(prev[mid]) < key) lo = mid + 1; else hi = mid; }
    bool matched = false;
    uint32_t from = 0xffffffffu;
    if (c.featureKey != 0)
        for (uint32_t t = lo; t < mp && pairKeyOf(prev[t]) == key; ++t)
            if (prev[t].featureKey == c.featureKey) {
                copyImpulses(c, prev[t]);
                if (c.vn0 < -1.0f && prev[t].vn0 < c.vn0) c.vn0 = prev[t].vn0;
                matched = true;
                from = t;
                break;
            }
    if (!matched) {
        if (c.vn0 < -1.0f)
            for (uint32_t t = lo; t < mp && pairKeyOf(prev[t]) == key; ++t)
                if (prev[t].vn0 < c.vn0) c.vn0 = prev[t].vn0;
        float best = 2.5e-3f;
        for (uint32_t t = lo; t < mp && pairKeyOf(prev[t]) == key; ++t) {
            float d = lengthSq(v3(prev[t].point) - v3(c.point));
            if (d < best) { best = d; copyImpulses(c, prev[t]); from = t; }
        }
    }
    cur[k] = c;
    if (match) match[k] = from == 0xffffffffu ? from : prevPos[from];
}

// The reference matches against its cache in the order std::sort leaves it
// (world.cpp:4518-4521), and std::sort is not stable: the first of several
// cached contacts with the contact's feature key, or the first of several at
// exactly the nearest distance, depends on that order.  Counts the contacts
// whose match such a tie decides (the cache here is in another order among a
// pair's contacts); the host then puts the cache in the exact order first.
// A tie whose candidates carry bit-identical warm-start data (the six
// impulses and vn0) decides nothing and is not counted, unless `anyTie` asks
// for every tie (the differentiable record keeps which entry was taken).
__device__ inline bool sameWarmData(const vx_contact& x, const vx_contact& y) {
    return __float_as_uint(x.normalImpulse) == __float_as_uint(y.normalImpulse) &&
           __float_as_uint(x.tangentImpulse1) == __float_as_uint(y.tangentImpulse1) &&
           __float_as_uint(x.tangentImpulse2) == __float_as_uint(y.tangentImpulse2) &&
           __float_as_uint(x.rollingImpulse1) == __float_as_uint(y.rollingImpulse1) &&
           __float_as_uint(x.rollingImpulse2) == __float_as_uint(y.rollingImpulse2) &&
           __float_as_uint(x.spinningImpulse) == __float_as_uint(y.spinningImpulse) &&
           __float_as_uint(x.vn0) == __float_as_uint(y.vn0);
}

__global__ __launch_bounds__(kThreads) void kWarmTies(const vx_contact* __restrict__ cur, uint32_t m,
                                                      const vx_contact* __restrict__ prev, uint32_t mp,
                                                      uint32_t* __restrict__ ties, bool anyTie) {
    uint32_t k = blockIdx.x * kThreads + threadIdx.x;
    bool tie = false;
    if (k < m && mp) {
        const vx_contact c = cur[k];
        const uint64_t key = pairKeyOf(c);
        uint32_t lo = 0, hi = mp;
        while (lo < hi) { uint32_t mid = (lo + hi) >> 1; if (pairKeyOf(prev[mid]) < key) lo = mid + 1; else hi = mid; }
        uint32_t same = 0, firstSame = 0;
        if (c.featureKey != 0)
            for (uint32_t t = lo; t < mp && pairKeyOf(prev[t]) == key; ++t)
                if (prev[t].featureKey == c.featureKey) {
                    if (same == 0) firstSame = t;
                    else if (anyTie || !sameWarmData(prev[firstSame], prev[t])) tie = true;
                    ++same;
                }
        if (same == 0) {
            float best = 2.5e-3f;
            uint32_t atBest = 0;
            bool differ = false;
            for (uint32_t t = lo; t < mp && pairKeyOf(prev[t]) == key; ++t) {
                const float d = lengthSq(v3(prev[t].point) - v3(c.point));
                if (d < best) { best = d; atBest = t; differ = false; }
                else if (d == best && best < 2.5e-3f && (anyTie || !sameWarmData(prev[atBest], prev[t]))) differ = true;
            }
            tie = differ;
        }
    }
    const uint64_t any = __ballot(tie);
    if (any && (threadIdx.x & 63) == 0) atomicAdd(ties, (uint32_t)__popcll(any));
}

// Cache contacts (canonical order) with their positions in the solve order
// they came from: (pair key, canonical index) at that position.
__global__ __launch_bounds__(kThreads) void kPrevSolveKeys(const vx_contact* __restrict__ prev, uint32_t mp,
                                                           const uint32_t* __restrict__ rank, int nb,
                                                           uint64_t* __restrict__ keys, uint32_t* __restrict__ vals) {
    uint32_t j = blockIdx.x * kThreads + threadIdx.x;
    if (j >= mp) return;
    const uint32_t r = rank[j];
    keys[r] = ((uint64_t)prev[j].a << nb) | prev[j].b;
    vals[r] = j;
}

// rank[order[i]] = i: the solve position of each canonical contact.
__global__ __launch_bounds__(kThreads) void kRankBack(const uint32_t* __restrict__ order, uint32_t m,
                                                      uint32_t* __restrict__ rank) {
    uint32_t i = blockIdx.x * kThreads + threadIdx.x;
    if (i < m) rank[order[i]] = i;
}

// out[t] = in[vals[t]].
__global__ __launch_bounds__(kThreads) void kGatherBy(const vx_contact* __restrict__ in, const uint32_t* __restrict__ vals,
                                                      uint32_t n, vx_contact* __restrict__ out) {
    uint32_t t = blockIdx.x * kThreads + threadIdx.x;
    if (t < n) out[t] = in[vals[t]];
}

// out[t] = in[vals[t]] for u32 words.
__global__ __launch_bounds__(kThreads) void kGatherU32(const uint32_t* __restrict__ in, const uint32_t* __restrict__ vals,
                                                       uint32_t n, uint32_t* __restrict__ out) {
    uint32_t t = blockIdx.x * kThreads + threadIdx.x;
    if (t < n) out[t] = in[vals[t]];
}

// World of each entry (by the first body of its pair key), for a stable
// world-major arrangement; per-world entry counts.
__global__ __launch_bounds__(kThreads) void kEntryWorlds(const uint64_t* __restrict__ keys, uint32_t n, int nb,
                                                         const uint32_t* __restrict__ worldOf,
                                                         uint64_t* __restrict__ wkeys, uint32_t* __restrict__ idx,
                                                         uint32_t* __restrict__ worldCount) {
    uint32_t i = blockIdx.x * kThreads + threadIdx.x;
    if (i >= n) return;
    const uint32_t w = worldOf[(uint32_t)(keys[i] >> nb)];
    wkeys[i] = w;
    idx[i] = i;
    atomicAdd(&worldCount[w], 1u);
}

__global__ __launch_bounds__(kThreads) void kPermuteKV(const uint64_t* __restrict__ keys, const uint32_t* __restrict__ vals,
                                                       const uint32_t* __restrict__ idx, uint32_t n,
                                                       uint64_t* __restrict__ keysOut, uint32_t* __restrict__ valsOut) {
    uint32_t i = blockIdx.x * kThreads + threadIdx.x;
    if (i >= n) return;
    keysOut[i] = keys[idx[i]];
    valsOut[i] = vals[idx[i]];
}

// --------------------------------------------------------------- coloring --
__device__ inline bool constrainedSide(const vx_body* bodies, const vx_contact& c, int side) {
    bool sensor = isSensor(bodies[c.a]) || isSensor(bodies[c.b]);
    return !sensor && isDynamic(bodies[side ? c.b : c.a]);
}

// One (body << cb | contact) entry per constrained contact side.
__global__ __launch_bounds__(kThreads) void kColorEntries(const vx_body* __restrict__ bodies, const vx_contact* __restrict__ cs,
                                                          uint32_t m, int cb, uint64_t* __restrict__ keys,
                                                          uint32_t* __restrict__ vals, uint32_t* __restrict__ counter) {
    uint32_t k = blockIdx.x * kThreads + threadIdx.x;
    bool s0 = false, s1 = false;
    vx_contact c;
    if (k < m) {
        c = cs[k];
        s0 = constrainedSide(bodies, c, 0);
        s1 = constrainedSide(bodies, c, 1);
    }
    uint32_t at = waveReserve((s0 ? 1u : 0u) + (s1 ? 1u : 0u), counter);
    if (s0) { keys[at] = ((uint64_t)c.a << cb) | k; vals[at] = k << 1; ++at; }
    if (s1) { keys[at] = ((uint64_t)c.b << cb) | k; vals[at] = (k << 1) | 1u; }
}

// For each constrained side, the previous contact side on the same body.
// (`eDev`: the entry count on the device, e then only bounds the grid.)
__global__ __launch_bounds__(kThreads) void kColorPreds(const uint64_t* __restrict__ keys, const uint32_t* __restrict__ vals,
                                                        uint32_t e, int cb, uint32_t* __restrict__ pred,
                                                        const uint32_t* __restrict__ eDev = nullptr) {
    uint32_t t = blockIdx.x * kThreads + threadIdx.x;
    if (eDev) e = *eDev;
    if (t >= e) return;
    uint32_t side = vals[t];
    uint32_t p = kNone;
    if (t > 0 && (keys[t - 1] >> cb) == (keys[t] >> cb)) p = vals[t - 1];
    pred[side] = p;
}

constexpr int kColorThreads = 1024;

// Greedy first-free-bit coloring (backend_cuda.cu:597-630) in contact order,
// reproduced exactly by dataflow: a contact's colour depends only on the
// masks its two bodies carry after their previous contacts, so each lane
// waits for those two predecessors and nothing else.  The lowest unfinished
// contact is always ready, hence progress; a spin budget still bounds every
// wait.  The kernels below differ in who takes which contacts and where the
// hand-offs live; they share the colour rule (greedyPick) and count colours
// in an LDS histogram flushed once per workgroup (one global atomic per
// colour and workgroup instead of one per contact).
constexpr uint32_t kColorHist = 256;  // colours counted in LDS; higher ones go to global atomics

__device__ inline int greedyPick(bool sa, bool sb, uint32_t pa, uint32_t pb, unsigned long long maskA,
                                 unsigned long long maskB, int& nextA, int& nextB, unsigned long long& outA,
                                 unsigned long long& outB) {
    if (sa || pa == kNone) { maskA = 0; nextA = 64; }
    if (sb || pb == kNone) { maskB = 0; nextB = 64; }
    const unsigned long long used = (sa ? 0ull : maskA) | (sb ? 0ull : maskB);
    int col;
    if (~used) {
        col = __ffsll((long long)~used) - 1;
    } else {
        int na = sa ? 64 : nextA, nb2 = sb ? 64 : nextB;
        col = na > nb2 ? na : nb2;
        if (!sa) nextA = col + 1;
        if (!sb) nextB = col + 1;
    }
    if (col < 64) {
        if (!sa) maskA |= 1ull << col;
        if (!sb) maskB |= 1ull << col;
    }
    outA = maskA;
    outB = maskB;
    return col;
}

struct ColorHist {
    uint32_t* hist;  // LDS, kColorHist entries
    uint32_t* top;   // LDS, largest colour + 1
    __device__ void init() {
        for (uint32_t i = threadIdx.x; i < kColorHist; i += blockDim.x) hist[i] = 0u;
        if (threadIdx.x == 0) *top = 0u;
        __syncthreads();
    }
    __device__ void add(int col, uint32_t* __restrict__ colorCount) {
        if ((uint32_t)col < kColorHist) atomicAdd(&hist[col], 1u);
        else atomicAdd(&colorCount[col], 1u);
        atomicMax(top, (uint32_t)col + 1u);
    }
    __device__ void flush(uint32_t* __restrict__ colorCount, uint32_t* __restrict__ numColors) {
        __syncthreads();
        for (uint32_t i = threadIdx.x; i < kColorHist; i += blockDim.x)
            if (hist[i]) atomicAdd(&colorCount[i], hist[i]);
        if (threadIdx.x == 0 && *top) atomicMax(numColors, *top);
    }
};

// One workgroup over all contacts; workgroup-scope release/acquire orders
// the hand-offs (all waves share one CU).
__global__ __launch_bounds__(kColorThreads) void kGreedyColor(const vx_body* __restrict__ bodies,
                                                              const vx_contact* __restrict__ cs, uint32_t m,
                                                              const uint32_t* __restrict__ pred,
                                                              unsigned long long* __restrict__ maskAfter,
                                                              int* __restrict__ nextAfter, uint32_t* color,
                                                              uint32_t* __restrict__ colorCount,
                                                              uint32_t* __restrict__ numColors, uint32_t* __restrict__ error) {
    __shared__ uint32_t hist[kColorHist], top;
    ColorHist h{hist, &top};
    h.init();
    uint32_t k = threadIdx.x;
    uint32_t spins = 0;
    while (k < m) {
        const uint32_t pa = pred[2 * k], pb = pred[2 * k + 1];
        const bool readyA = pa == kNone || __hip_atomic_load(&color[pa >> 1], __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_WORKGROUP) != kNone;
        const bool readyB = pb == kNone || __hip_atomic_load(&color[pb >> 1], __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_WORKGROUP) != kNone;
        if (readyA && readyB) {
            const vx_contact& c = cs[k];
            const bool sa = !constrainedSide(bodies, c, 0), sb = !constrainedSide(bodies, c, 1);
            int nextA = 0, nextB = 0;
            unsigned long long mA = 0, mB = 0, maskA = 0, maskB = 0;
            if (!sa && pa != kNone) { mA = maskAfter[pa]; nextA = nextAfter[pa]; }
            if (!sb && pb != kNone) { mB = maskAfter[pb]; nextB = nextAfter[pb]; }
            const int col = greedyPick(sa, sb, pa, pb, mA, mB, nextA, nextB, maskA, maskB);
            maskAfter[2 * k] = maskA; nextAfter[2 * k] = nextA;
            maskAfter[2 * k + 1] = maskB; nextAfter[2 * k + 1] = nextB;
            h.add(col, colorCount);
            __hip_atomic_store(&color[k], (uint32_t)col, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_WORKGROUP);
            k += kColorThreads;
            spins = 0;
        } else if (++spins > (1u << 26)) {
            atomicOr(error, 1u);
            break;
        }
    }
    h.flush(colorCount, numColors);
}

// Batched worlds whose bodies lie world after world: the contacts of one
// world are contiguous in contact order and none has a predecessor outside
// it, so each world is coloured by one workgroup over its range
// [segStart[w], segStart[w + 1]), every world independently.  A world of up
// to kWorldColorLds contacts keeps its colours and masks in LDS (hand-offs
// through LDS, results copied out at the end); a larger one works on the
// global arrays as kGreedyColor does.  A lane reads its contact's
// predecessors and constrained-side flags before it waits.
constexpr uint32_t kWorldColorLds = 2048;

__global__ __launch_bounds__(kThreads) void kGreedyColorWorlds(const vx_body* __restrict__ bodies,
                                                               const vx_contact* __restrict__ cs,
                                                               const uint32_t* __restrict__ segStart,
                                                               const uint32_t* __restrict__ pred,
                                                               unsigned long long* __restrict__ maskAfter,
                                                               int* __restrict__ nextAfter, uint32_t* color,
                                                               uint32_t* __restrict__ colorCount,
                                                               uint32_t* __restrict__ numColors,
                                                               uint32_t* __restrict__ error) {
    __shared__ uint32_t hist[kColorHist], top;
    __shared__ uint32_t lcol[kWorldColorLds];
    __shared__ unsigned long long lmask[2 * kWorldColorLds];
    __shared__ int lnext[2 * kWorldColorLds];
    ColorHist h{hist, &top};
    const uint32_t begin = segStart[blockIdx.x], end = segStart[blockIdx.x + 1];
    const bool inLds = end - begin <= kWorldColorLds;
    for (uint32_t i = threadIdx.x; inLds && i < end - begin; i += kThreads) lcol[i] = kNone;
    h.init();
    uint32_t k = begin + threadIdx.x;
    uint32_t spins = 0;
    bool loaded = false, sa = false, sb = false;
    uint32_t pa = kNone, pb = kNone;
    while (k < end) {
        if (!loaded) {
            pa = pred[2 * k];
            pb = pred[2 * k + 1];
            const vx_contact& c = cs[k];
            sa = !constrainedSide(bodies, c, 0);
            sb = !constrainedSide(bodies, c, 1);
            loaded = true;
        }
        bool readyA, readyB;
        if (inLds) {
            readyA = pa == kNone || __hip_atomic_load(&lcol[(pa >> 1) - begin], __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_WORKGROUP) != kNone;
            readyB = pb == kNone || __hip_atomic_load(&lcol[(pb >> 1) - begin], __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_WORKGROUP) != kNone;
        } else {
            readyA = pa == kNone || __hip_atomic_load(&color[pa >> 1], __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_WORKGROUP) != kNone;
            readyB = pb == kNone || __hip_atomic_load(&color[pb >> 1], __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_WORKGROUP) != kNone;
        }
        if (readyA && readyB) {
            int nextA = 0, nextB = 0;
            unsigned long long mA = 0, mB = 0, maskA = 0, maskB = 0;
            if (inLds) {
                if (!sa && pa != kNone) { mA = lmask[pa - 2 * begin]; nextA = lnext[pa - 2 * begin]; }
                if (!sb && pb != kNone) { mB = lmask[pb - 2 * begin]; nextB = lnext[pb - 2 * begin]; }
            } else {
                if (!sa && pa != kNone) { mA = maskAfter[pa]; nextA = nextAfter[pa]; }
                if (!sb && pb != kNone) { mB = maskAfter[pb]; nextB = nextAfter[pb]; }
            }
            const int col = greedyPick(sa, sb, pa, pb, mA, mB, nextA, nextB, maskA, maskB);
            if (inLds) {
                const uint32_t j = 2 * (k - begin);
                lmask[j] = maskA; lnext[j] = nextA;
                lmask[j + 1] = maskB; lnext[j + 1] = nextB;
                color[k] = (uint32_t)col;
                __hip_atomic_store(&lcol[k - begin], (uint32_t)col, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_WORKGROUP);
            } else {
                maskAfter[2 * k] = maskA; nextAfter[2 * k] = nextA;
                maskAfter[2 * k + 1] = maskB; nextAfter[2 * k + 1] = nextB;
                __hip_atomic_store(&color[k], (uint32_t)col, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_WORKGROUP);
            }
            h.add(col, colorCount);
            k += kThreads;
            loaded = false;
            spins = 0;
        } else if (++spins > (1u << 26)) {
            atomicOr(error, 1u);
            break;
        }
    }
    h.flush(colorCount, numColors);
}

// Contacts per world (by the first body), for kGreedyColorWorlds' ranges.
__global__ __launch_bounds__(kThreads) void kContactWorlds(const vx_contact* __restrict__ cs, uint32_t m,
                                                           const uint32_t* __restrict__ worldOf,
                                                           uint32_t* __restrict__ count) {
    const uint32_t k = blockIdx.x * kThreads + threadIdx.x;
    if (k < m) atomicAdd(&count[worldOf[cs[k].a]], 1u);
}

// The same greedy colouring by dataflow across the whole device, for large
// contact sets: lanes walk contacts with a grid-wide stride, wait for their
// predecessors with relaxed device-scope loads, and take one device-scope
// acquire before reading the predecessors' masks; a colour is published with
// a device-scope release.  The grid is sized to be co-resident, so the lowest
// uncoloured contact is always running and ready.
constexpr int kColorBlocks = 256;

__global__ __launch_bounds__(kThreads) void kGreedyColorDevice(const vx_body* __restrict__ bodies,
                                                               const vx_contact* __restrict__ cs, uint32_t m,
                                                               const uint32_t* __restrict__ pred,
                                                               unsigned long long* __restrict__ maskAfter,
                                                               int* __restrict__ nextAfter, uint32_t* color,
                                                               uint32_t* __restrict__ colorCount,
                                                               uint32_t* __restrict__ numColors,
                                                               uint32_t* __restrict__ error) {
    __shared__ uint32_t hist[kColorHist], top;
    ColorHist h{hist, &top};
    h.init();
    const uint32_t stride = gridDim.x * kThreads;
    uint32_t k = blockIdx.x * kThreads + threadIdx.x;
    uint32_t spins = 0;
    while (k < m) {
        const uint32_t pa = pred[2 * k], pb = pred[2 * k + 1];
        const bool readyA = pa == kNone || __hip_atomic_load(&color[pa >> 1], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) != kNone;
        const bool readyB = pb == kNone || __hip_atomic_load(&color[pb >> 1], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) != kNone;
        if (readyA && readyB) {
            if (pa != kNone || pb != kNone) __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
            const vx_contact& c = cs[k];
            const bool sa = !constrainedSide(bodies, c, 0), sb = !constrainedSide(bodies, c, 1);
            int nextA = 0, nextB = 0;
            unsigned long long mA = 0, mB = 0, maskA = 0, maskB = 0;
            if (!sa && pa != kNone) { mA = maskAfter[pa]; nextA = nextAfter[pa]; }
            if (!sb && pb != kNone) { mB = maskAfter[pb]; nextB = nextAfter[pb]; }
            const int col = greedyPick(sa, sb, pa, pb, mA, mB, nextA, nextB, maskA, maskB);
            maskAfter[2 * k] = maskA; nextAfter[2 * k] = nextA;
            maskAfter[2 * k + 1] = maskB; nextAfter[2 * k + 1] = nextB;
            h.add(col, colorCount);
            __hip_atomic_store(&color[k], (uint32_t)col, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
            k += stride;
            spins = 0;
        } else if (++spins > (1u << 26)) {
            atomicOr(error, 1u);
            break;
        }
    }
    h.flush(colorCount, numColors);
}

// ----------------------------------------------------------------- solver --
__global__ __launch_bounds__(kThreads) void kGather(const vx_contact* __restrict__ in, const uint32_t* __restrict__ order,
                                                    uint32_t m, vx_contact* __restrict__ out) {
    uint32_t k = blockIdx.x * kThreads + threadIdx.x;
    if (k < m) out[k] = in[order[k]];
}

__global__ __launch_bounds__(kThreads) void kScatterBack(const vx_contact* __restrict__ in, const uint32_t* __restrict__ order,
                                                         uint32_t m, vx_contact* __restrict__ out) {
    uint32_t k = blockIdx.x * kThreads + threadIdx.x;
    if (k < m) out[order[k]] = in[k];
}


// ------------------------------------------------------------ contact events --
// world.cpp:4399-4485.  Contacts in solve order (as the solver left them) are
// keyed by their canonical pair (lo << nb | hi); inactive contacts get the
// key `inactive` so that a stable sort puts them last and keeps, within a
// pair, the solve order that picks the representative on ties.
__global__ __launch_bounds__(kThreads) void kEventKeys(const vx_body* __restrict__ bodies,
                                                       const vx_contact* __restrict__ cs, uint32_t m, float dt, int nb,
                                                       uint64_t inactive, uint64_t* __restrict__ keys,
                                                       uint32_t* __restrict__ vals, uint32_t* __restrict__ activeCount) {
    uint32_t k = blockIdx.x * kThreads + threadIdx.x;
    bool act = false;
    if (k < m) {
        const vx_contact& c = cs[k];
        const vx_body &a = bodies[c.a], &b = bodies[c.b];
        const bool sensor = isSensor(a) || isSensor(b);
        const float lg = liveGap(a, b, c);
        const bool swept = sensor && c.vn0 < 0.0f && c.gap <= -c.vn0 * dt + 1e-3f;
        act = !(c.normalImpulse <= 1e-6f && lg > 1e-3f && !swept);
        const uint32_t lo = c.a < c.b ? c.a : c.b, hi = c.a < c.b ? c.b : c.a;
        keys[k] = act ? ((uint64_t)lo << nb) | hi : inactive;
        vals[k] = k;
    }
    waveReserve(act ? 1u : 0u, activeCount);
}

__device__ inline uint32_t lowerBound64(const uint64_t* keys, uint32_t n, uint64_t key) {
    uint32_t lo = 0, hi = n;
    while (lo < hi) {
        const uint32_t mid = (lo + hi) >> 1;
        if (keys[mid] < key) lo = mid + 1; else hi = mid;
    }
    return lo;
}

__device__ inline vx_contact_event eventOf(const vx_body* bodies, uint32_t lo, uint32_t hi, uint32_t type) {
    vx_contact_event e;
    e.a = lo; e.b = hi;
    e.point[0] = e.point[1] = e.point[2] = 0.0f;
    e.normal[0] = e.normal[1] = e.normal[2] = 0.0f;
    e.impulse = 0.0f;
    e.type = type;
    e.sensor = (isSensor(bodies[lo]) || isSensor(bodies[hi])) ? 1u : 0u;
    return e;
}

// Representative contact of a pair: the narrow-phase normal points from B to
// A, events from the high body to the low one (world.cpp:4443-4452).
__device__ inline void eventRep(vx_contact_event& e, const vx_contact& c, uint32_t lo) {
    e.impulse = c.normalImpulse;
    e.point[0] = c.point.x; e.point[1] = c.point.y; e.point[2] = c.point.z;
    e.normal[0] = c.a == lo ? -c.normal.x : c.normal.x;
    e.normal[1] = c.a == lo ? -c.normal.y : c.normal.y;
    e.normal[2] = c.a == lo ? -c.normal.z : c.normal.z;
}

__global__ __launch_bounds__(kThreads) void kEventRunHeads(const uint64_t* __restrict__ keys,
                                                           const uint32_t* __restrict__ activeCount, uint32_t m,
                                                           uint32_t* __restrict__ head) {
    uint32_t t = blockIdx.x * kThreads + threadIdx.x;
    if (t >= m) return;
    head[t] = t < *activeCount && (t == 0 || keys[t - 1] != keys[t]) ? 1u : 0u;
}

// One event per active pair, in pair-key order: Begin, or Persist when the
// pair was active in the previous step; the representative is the contact
// with the largest normal impulse (the first in solve order on ties).
__global__ __launch_bounds__(kThreads) void kEventRuns(const vx_body* __restrict__ bodies,
                                                       const vx_contact* __restrict__ cs,
                                                       const uint64_t* __restrict__ keys,
                                                       const uint32_t* __restrict__ vals,
                                                       const uint32_t* __restrict__ head,
                                                       const uint32_t* __restrict__ runPos,
                                                       const uint32_t* __restrict__ activeCount, uint32_t m, int nb,
                                                       const uint64_t* __restrict__ prevKeys, uint32_t np,
                                                       vx_contact_event* __restrict__ events,
                                                       uint64_t* __restrict__ curKeys) {
    uint32_t t = blockIdx.x * kThreads + threadIdx.x;
    if (t >= m || !head[t]) return;
    const uint32_t na = *activeCount;
    const uint64_t key = keys[t];
    const uint32_t lo = (uint32_t)(key >> nb), hi = (uint32_t)(key & ((1ull << nb) - 1));
    const uint32_t at = lowerBound64(prevKeys, np, key);
    vx_contact_event e = eventOf(bodies, lo, hi, at < np && prevKeys[at] == key ? 1u : 0u);
    bool set = false;
    for (uint32_t u = t; u < na && keys[u] == key; ++u) {
        const vx_contact& c = cs[vals[u]];
        if (!set || c.normalImpulse > e.impulse) { set = true; eventRep(e, c, lo); }
    }
    events[runPos[t]] = e;
    curKeys[runPos[t]] = key;
}

// Pairs of the previous step (prevKeys: ascending, unique) missing from this
// step's active keys A = curKeys[0, nA): End events, or Persist when both
// bodies are inert (static or fully asleep), after the active pairs in
// previous-key order (world.cpp:4454-4481).
//
// The reference looks each previous key up with std::binary_search over a
// list that grows as inert keys are appended behind A.  Previous keys ascend,
// so every appended key is below the key searched, and the search sees the
// predicate "element < key" as T^p F^q T^b: the p keys of A below the key,
// the q = nA - p others, then the b keys appended so far.  libstdc++'s
// lower_bound over that pattern ends at p or at the end (checked for all
// p, q, b < 40), so a key absent from A is never found, and a key of A is
// found only if the search over (p, q, b) ends at p.  A pair in A is never
// inert (the broad phase drops pairs of two inert bodies), so the appended
// keys are the inert absent ones, and b follows from a scan.
__device__ inline bool growingSearchFinds(uint32_t p, uint32_t nA, uint32_t b) {
    uint32_t first = 0, len = nA + b;
    while (len > 0) {
        const uint32_t half = len >> 1, mid = first + half;
        if (mid < p || mid >= nA) { first = mid + 1; len -= half + 1; } else len = half;
    }
    return first == p;
}

enum : uint32_t { kEndInA = 1u, kEndValid = 2u, kEndInert = 4u };

// Per previous key: its lower bound p in A, whether A holds it, whether both
// bodies exist and are inert; appended[i] = 1 for an inert key absent from A.
__global__ __launch_bounds__(kThreads) void kEventEndsClassify(const vx_body* __restrict__ bodies, uint32_t n, int nb,
                                                               const uint64_t* __restrict__ prevKeys, uint32_t np,
                                                               const uint64_t* __restrict__ curKeys,
                                                               const uint32_t* __restrict__ runs,
                                                               uint32_t* __restrict__ endP, uint32_t* __restrict__ endBits,
                                                               uint32_t* __restrict__ appended) {
    uint32_t i = blockIdx.x * kThreads + threadIdx.x;
    if (i >= np) return;
    const uint64_t key = prevKeys[i];
    const uint32_t nA = *runs;
    const uint32_t p = lowerBound64(curKeys, nA, key);
    const uint32_t lo = (uint32_t)(key >> nb), hi = (uint32_t)(key & ((1ull << nb) - 1));
    uint32_t bits = p < nA && curKeys[p] == key ? kEndInA : 0u;
    if (lo < n && hi < n) {
        bits |= kEndValid;
        if (inert(bodies[lo]) && inert(bodies[hi])) bits |= kEndInert;
    }
    endP[i] = p;
    endBits[i] = bits;
    appended[i] = !(bits & kEndInA) && (bits & kEndInert) ? 1u : 0u;
}

__device__ inline bool endEmits(uint32_t bits, uint32_t p, uint32_t nA, uint32_t b) {
    const bool found = (bits & kEndInA) && growingSearchFinds(p, nA, b);
    return !found && (bits & kEndValid);
}

// emit[i] = 1 for a key that yields an event; appended keys are compacted
// (ascending) into appKeys at their rank.
__global__ __launch_bounds__(kThreads) void kEventEndsEmit(const uint64_t* __restrict__ prevKeys, uint32_t np,
                                                           const uint32_t* __restrict__ runs,
                                                           const uint32_t* __restrict__ endP,
                                                           const uint32_t* __restrict__ endBits,
                                                           const uint32_t* __restrict__ appRank,
                                                           uint32_t* __restrict__ emit, uint64_t* __restrict__ appKeys) {
    uint32_t i = blockIdx.x * kThreads + threadIdx.x;
    if (i >= np) return;
    const uint32_t bits = endBits[i], b = appRank[i];
    emit[i] = endEmits(bits, endP[i], *runs, b) ? 1u : 0u;
    if (!(bits & kEndInA) && (bits & kEndInert)) appKeys[b] = prevKeys[i];
}

// First contact of pair {lo, hi} in the previous step's cache, which is sorted
// by (a << 32 | b); a pair's contacts share one orientation (a, b), either
// way round, and the (lo, hi) run sorts before a (hi, lo) run.  mp when the
// pair has none.
__device__ inline uint32_t cachedRun(const vx_contact* prev, uint32_t mp, uint64_t key) {
    uint32_t l = 0, h = mp;
    while (l < h) { const uint32_t mid = (l + h) >> 1; if (pairKeyOf(prev[mid]) < key) l = mid + 1; else h = mid; }
    return l < mp && pairKeyOf(prev[l]) == key ? l : mp;
}

__device__ inline uint32_t firstCached(const vx_contact* prev, uint32_t mp, uint32_t lo, uint32_t hi) {
    const uint32_t t = cachedRun(prev, mp, (uint64_t)lo << 32 | hi);
    return t < mp ? t : cachedRun(prev, mp, (uint64_t)hi << 32 | lo);
}

// The events: Persist (representative: the pair's first cached contact) for
// inert pairs, End otherwise.
__global__ __launch_bounds__(kThreads) void kEventEndsWrite(const vx_body* __restrict__ bodies, int nb,
                                                            const uint64_t* __restrict__ prevKeys, uint32_t np,
                                                            const uint32_t* __restrict__ runs,
                                                            const uint32_t* __restrict__ endP,
                                                            const uint32_t* __restrict__ endBits,
                                                            const uint32_t* __restrict__ appRank,
                                                            const uint32_t* __restrict__ emitPos,
                                                            const vx_contact* __restrict__ prev, uint32_t mp,
                                                            vx_contact_event* __restrict__ events) {
    uint32_t i = blockIdx.x * kThreads + threadIdx.x;
    if (i >= np) return;
    const uint32_t bits = endBits[i], nA = *runs;
    if (!endEmits(bits, endP[i], nA, appRank[i])) return;
    const uint64_t key = prevKeys[i];
    const uint32_t lo = (uint32_t)(key >> nb), hi = (uint32_t)(key & ((1ull << nb) - 1));
    const bool persist = !(bits & kEndInA) && (bits & kEndInert);
    vx_contact_event e = eventOf(bodies, lo, hi, persist ? 1u : 2u);
    if (persist) {
        const uint32_t t = firstCached(prev, mp, lo, hi);
        if (t < mp) eventRep(e, prev[t], lo);
    }
    events[nA + emitPos[i]] = e;
}

// The key list handed to the next step: A merged with the appended keys
// (disjoint, both ascending).
__global__ __launch_bounds__(kThreads) void kEventKeysMerge(const uint64_t* __restrict__ curKeys,
                                                            const uint32_t* __restrict__ runs,
                                                            const uint64_t* __restrict__ appKeys,
                                                            const uint32_t* __restrict__ appCount, uint32_t cap,
                                                            uint64_t* __restrict__ merged) {
    uint32_t j = blockIdx.x * kThreads + threadIdx.x;
    const uint32_t nA = *runs, nB = *appCount;
    if (j >= cap || j >= nA + nB) return;
    if (j < nA) merged[j + lowerBound64(appKeys, nB, curKeys[j])] = curKeys[j];
    else merged[(j - nA) + lowerBound64(curKeys, nA, appKeys[j - nA])] = appKeys[j - nA];
}

// ----------------------------------------------------- warm-start cache --
// Contacts cached for the next step (world.cpp:4492-4521): this step's, plus,
// for every pair an inert Persist event kept alive, the first of its cached
// contacts (later ones of that pair count as already cached).
__global__ __launch_bounds__(kThreads) void kCarryFlags(const vx_contact* __restrict__ prev, uint32_t mp, int nb,
                                                        const uint64_t* __restrict__ appKeys,
                                                        const uint32_t* __restrict__ appCount,
                                                        uint32_t* __restrict__ carry) {
    uint32_t t = blockIdx.x * kThreads + threadIdx.x;
    if (t >= mp) return;
    const vx_contact& c = prev[t];
    bool keep = t == 0 || pairKeyOf(prev[t - 1]) != pairKeyOf(c);
    if (keep) {
        const uint32_t lo = c.a < c.b ? c.a : c.b, hi = c.a < c.b ? c.b : c.a;
        const uint64_t key = ((uint64_t)lo << nb) | hi;
        const uint32_t nB = *appCount, at = lowerBound64(appKeys, nB, key);
        keep = at < nB && appKeys[at] == key;
    }
    carry[t] = keep ? 1u : 0u;
}

// The cache candidates in the order the reference hands them to std::sort
// (world.cpp:4496-4521): this step's contacts in solve order, then the
// carried ones in cache order; (pair key, source) with sources below m
// naming canonical contacts of this step and m + t the cached contact t.
__global__ __launch_bounds__(kThreads) void kCacheSolveKeys(const vx_contact* __restrict__ solve,
                                                            const uint32_t* __restrict__ order, uint32_t m,
                                                            const vx_contact* __restrict__ prev, uint32_t mp, int nb,
                                                            const uint32_t* __restrict__ carry,
                                                            const uint32_t* __restrict__ carryRank,
                                                            uint64_t* __restrict__ keys, uint32_t* __restrict__ vals) {
    uint32_t j = blockIdx.x * kThreads + threadIdx.x;
    if (j < m) {
        keys[j] = ((uint64_t)solve[j].a << nb) | solve[j].b;
        vals[j] = order[j];
    } else if (j < m + mp && carry[j - m]) {
        const uint32_t t = j - m, r = m + carryRank[t];
        keys[r] = ((uint64_t)prev[t].a << nb) | prev[t].b;
        vals[r] = m + t;
    }
}

__global__ __launch_bounds__(kThreads) void kCacheGather(const vx_contact* __restrict__ cur, uint32_t m,
                                                         const vx_contact* __restrict__ prev, uint32_t mp,
                                                         const uint32_t* __restrict__ vals,
                                                         const uint32_t* __restrict__ carried,
                                                         vx_contact* __restrict__ out) {
    uint32_t t = blockIdx.x * kThreads + threadIdx.x;
    if (t >= m + mp || t >= m + *carried) return;
    const uint32_t v = vals[t];
    out[t] = v < m ? cur[v] : prev[v - m];
}

// --------------------------------------------------------- origin shift --
// World::shiftOrigin (world.cpp:1978-2036): every stored world coordinate
// moves by -offset.  Run once with apply = false to flag a result that is not
// finite (the reference checks everything before it changes anything), then
// with apply = true.
__device__ inline bool finite3(V3 v) { return isfinite(v.x) && isfinite(v.y) && isfinite(v.z); }

__device__ inline void shiftVec(vx_vec3& p, V3 off, bool apply, bool& bad) {
    const V3 r = v3(p) - off;
    if (!finite3(r)) bad = true;
    if (apply) p = raw(r);
}

__global__ __launch_bounds__(kThreads) void kShiftBodies(vx_body* __restrict__ bodies, uint32_t n, V3 off, int apply,
                                                         uint32_t* __restrict__ bad) {
    uint32_t i = blockIdx.x * kThreads + threadIdx.x;
    if (i >= n) return;
    vx_body& b = bodies[i];
    bool e = false;
    shiftVec(b.position, off, apply, e);
    if (b.shape == VX_SHAPE_PLANE) {
        const float po = b.planeOffset - dot(v3(b.planeNormal), off);
        if (!isfinite(po)) e = true;
        if (apply) b.planeOffset = po;
    }
    if (e) atomicOr(bad, 1u);
}

__global__ __launch_bounds__(kThreads) void kShiftPoints(vx_vec3* __restrict__ pts, uint32_t count, V3 off, int apply,
                                                         uint32_t* __restrict__ bad) {
    uint32_t i = blockIdx.x * kThreads + threadIdx.x;
    if (i >= count) return;
    bool e = false;
    shiftVec(pts[i], off, apply, e);
    if (e) atomicOr(bad, 1u);
}

template <class T>  // vx_mesh, vx_bvh_node: aabbMin / aabbMax
__global__ __launch_bounds__(kThreads) void kShiftBounds(T* __restrict__ items, uint32_t count, V3 off, int apply,
                                                         uint32_t* __restrict__ bad) {
    uint32_t i = blockIdx.x * kThreads + threadIdx.x;
    if (i >= count) return;
    bool e = false;
    shiftVec(items[i].aabbMin, off, apply, e);
    shiftVec(items[i].aabbMax, off, apply, e);
    if (e) atomicOr(bad, 1u);
}

__global__ __launch_bounds__(kThreads) void kShiftContacts(vx_contact* __restrict__ cs, uint32_t count, V3 off,
                                                           int apply, uint32_t* __restrict__ bad) {
    uint32_t i = blockIdx.x * kThreads + threadIdx.x;
    if (i >= count) return;
    bool e = false;
    shiftVec(cs[i].point, off, apply, e);
    if (e) atomicOr(bad, 1u);
}

__global__ __launch_bounds__(kThreads) void kShiftEvents(vx_contact_event* __restrict__ ev, uint32_t count, V3 off,
                                                         int apply, uint32_t* __restrict__ bad) {
    uint32_t i = blockIdx.x * kThreads + threadIdx.x;
    if (i >= count || ev[i].type == 2u) return;  // End events carry no point
    const V3 r = V3{ev[i].point[0], ev[i].point[1], ev[i].point[2]} - off;
    if (!finite3(r)) atomicOr(bad, 1u);
    if (apply) { ev[i].point[0] = r.x; ev[i].point[1] = r.y; ev[i].point[2] = r.z; }
}

// world.cpp:4364-4397, one colour at a time.
__device__ __attribute__((always_inline)) inline void positionRow(vx_body* __restrict__ bodies,
                                                                 const vx_contact* __restrict__ cs, uint32_t k,
                                                                 float h, const vx_solver_options& opt) {
    const vx_contact c = cs[k];
    vx_body a = bodies[c.a], b = bodies[c.b];
    if (isSensor(a) || isSensor(b)) return;
    float ims = solverInvMass(a) + solverInvMass(b);
    if (ims <= 0.0f) return;
    float g = liveGap(a, b, c);
    if (g >= -1e-3f) return;
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
    if (isDynamic(a) && !asleepFully(a)) bodies[c.a].position = raw(v3(a.position) + v3(c.normal) * (push * solverInvMass(a)));
    if (isDynamic(b) && !asleepFully(b)) bodies[c.b].position = raw(v3(b.position) - v3(c.normal) * (push * solverInvMass(b)));
}

__global__ __launch_bounds__(kThreads) void kPositionColor(vx_body* __restrict__ bodies, const vx_contact* __restrict__ cs,
                                                           const uint32_t* __restrict__ colorStart, uint32_t color, float h,
                                                           vx_solver_options opt) {
    uint32_t first = colorStart[color], count = colorStart[color + 1] - first;
    uint32_t k = blockIdx.x * kThreads + threadIdx.x;
    if (k >= count) return;
    positionRow(bodies, cs, first + k, h, opt);
}

// A run of narrow colours [c0, c1) in one workgroup, one colour after the
// other: the bodies a colour writes are read by later colours only after the
// workgroup barrier, on the same compute unit.
constexpr int kRunThreads = 1024;
__global__ __launch_bounds__(kRunThreads) void kPositionColorRun(vx_body* __restrict__ bodies,
                                                                 const vx_contact* __restrict__ cs,
                                                                 const uint32_t* __restrict__ colorStart, uint32_t c0,
                                                                 uint32_t c1, float h, vx_solver_options opt) {
    for (uint32_t c = c0; c < c1; ++c) {
        const uint32_t first = colorStart[c], last = colorStart[c + 1];
        for (uint32_t k = first + threadIdx.x; k < last; k += kRunThreads) positionRow(bodies, cs, k, h, opt);
        __syncthreads();
    }
}

}  // namespace vxk
