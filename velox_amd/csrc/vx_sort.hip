// Hand-written device sort and scan primitives for the HIP device step.
//
//  * LSD radix sort of (u64 key, u32 value) pairs over the low `bits` of the
//    key, 8-bit digits, one launch per pass (tile offsets by decoupled
//    look-back) after one launch that counts every pass's digits.  The pass's
//    scatter ranks every lane among equal digits of its wavefront with eight
//    ballots (a 64-lane match mask) and chains waves through LDS.  Stability
//    per pass is what makes the multi-pass LSD sort correct.
//  * Exclusive scan of u32 counts: one workgroup up to 16384 counts, else
//    tile scans and an add-back whose workgroups sum the earlier tiles'
//    totals themselves.
#include <algorithm>
#include "vx_internal.hpp"
#include "vx_stdsort.hpp"

namespace vxk {

constexpr int kSortThreads = 256;          // 4 wavefronts
constexpr int kSortWaves = kSortThreads / 64;
constexpr int kSortItems = 4;              // rounds per tile
constexpr int kSortTile = kSortThreads * kSortItems;

__device__ inline uint32_t digitOf(uint64_t key, int shift) { return (uint32_t)(key >> shift) & 0xffu; }

// One launch per pass.  Before the passes, one read of the keys counts the
// digits of every pass (kRadixDigitCounts).  A pass's workgroup takes the
// next tile from a counter (so every lower tile has started), counts its
// tile's digits, publishes them for its successors, and finds the tile's
// offset per digit by looking back over its predecessors' published words
// (decoupled look-back: a word holds either the tile's own count, or the
// inclusive prefix through that tile; the walk adds own counts until it
// meets a prefix).  Tile 0 publishes its prefix at once, so the walk always
// ends, and a tile waits only on tiles that started before it.  The scatter
// then ranks each lane among equal digits of its wavefront with eight
// ballots and chains the wavefronts through LDS, so a pass is stable.
constexpr uint32_t kLookAggregate = 1u << 30, kLookPrefix = 2u << 30, kLookCount = (1u << 30) - 1;
constexpr int kMaxPasses = 8;

// `nDev`: the element count on the device (then `n` is only the capacity the
// grid was sized for).
// With `keysCopy` it also copies (keys, vals) there: a sort of an odd number
// of passes then runs from the copy, so its last pass writes the input
// buffers and needs no copy back.
__global__ __launch_bounds__(kSortThreads) void kRadixDigitCounts(const uint64_t* __restrict__ keys,
                                                                  const uint32_t* __restrict__ vals, uint32_t n,
                                                                  const uint32_t* __restrict__ nDev, int passes,
                                                                  uint32_t* __restrict__ digitCounts,
                                                                  uint64_t* __restrict__ keysCopy,
                                                                  uint32_t* __restrict__ valsCopy) {
    __shared__ uint32_t counts[kMaxPasses][256];
    if (nDev) n = *nDev;
    if (blockIdx.x * kSortTile >= n) return;
    for (int p = 0; p < kMaxPasses; ++p) counts[p][threadIdx.x] = 0;
    __syncthreads();
    const uint32_t base = blockIdx.x * kSortTile;
    // The tile's keys are loaded together (one memory round trip), then counted.
    uint64_t kr[kSortItems];
#pragma unroll
    for (int r = 0; r < kSortItems; ++r) {
        const uint32_t i = base + r * kSortThreads + threadIdx.x;
        kr[r] = i < n ? keys[i] : 0;
    }
#pragma unroll
    for (int r = 0; r < kSortItems; ++r) {
        const uint32_t i = base + r * kSortThreads + threadIdx.x;
        if (i < n) {
            for (int p = 0; p < passes; ++p) atomicAdd(&counts[p][digitOf(kr[r], 8 * p)], 1u);
            if (keysCopy) {
                keysCopy[i] = kr[r];
                valsCopy[i] = vals[i];
            }
        }
    }
    __syncthreads();
    for (int p = 0; p < passes; ++p)
        if (counts[p][threadIdx.x]) atomicAdd(&digitCounts[p * 256 + threadIdx.x], counts[p][threadIdx.x]);
}

__global__ __launch_bounds__(kSortThreads) void kRadixPass(const uint64_t* __restrict__ keysIn,
                                                           const uint32_t* __restrict__ valsIn, uint32_t n,
                                                           const uint32_t* __restrict__ nDev, int shift,
                                                           const uint32_t* __restrict__ digitCounts,
                                                           uint32_t* __restrict__ tileCounter, uint32_t* status,
                                                           uint64_t* __restrict__ keysOut,
                                                           uint32_t* __restrict__ valsOut) {
    __shared__ uint32_t running[256];
    __shared__ uint32_t waveCount[kSortWaves][256];
    __shared__ uint32_t mine[256];
    __shared__ uint32_t waveSums[kSortWaves];
    __shared__ uint32_t tileS;
    const uint32_t d = threadIdx.x;
    if (nDev) n = *nDev;
    if (d == 0) tileS = atomicAdd(tileCounter, 1u);
    mine[d] = 0;
    for (int w = 0; w < kSortWaves; ++w) waveCount[w][d] = 0;
    __syncthreads();
    const uint32_t tile = tileS;
    const uint32_t base = tile * kSortTile;
    if (base >= n) return;  // past the last tile: no tile waits on it
    // The tile's keys and values stay in registers from one load (issued
    // together) to the scatter.
    uint64_t kr[kSortItems];
    uint32_t vr[kSortItems];
#pragma unroll
    for (int r = 0; r < kSortItems; ++r) {
        const uint32_t i = base + r * kSortThreads + threadIdx.x;
        kr[r] = i < n ? keysIn[i] : 0;
        vr[r] = i < n ? valsIn[i] : 0;
    }
#pragma unroll
    for (int r = 0; r < kSortItems; ++r) {
        const uint32_t i = base + r * kSortThreads + threadIdx.x;
        if (i < n) atomicAdd(&mine[digitOf(kr[r], shift)], 1u);
    }
    __syncthreads();
    const uint32_t own = mine[d];
    uint32_t* word = status + (size_t)tile * 256 + d;
    __hip_atomic_store(word, (tile ? kLookAggregate : kLookPrefix) | own, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    // The digit's first position in the output: exclusive scan of the pass's
    // digit counts over the 256 digits.
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const uint32_t c = digitCounts[d];
    uint32_t x = c;
#pragma unroll
    for (int off = 1; off < 64; off <<= 1) {
        const uint32_t y = __shfl_up(x, off, 64);
        if (lane >= off) x += y;
    }
    if (lane == 63) waveSums[wave] = x;
    // Look-back: the counts of this digit in all earlier tiles.  The words of
    // kLookBatch predecessors are loaded together and consumed in tile order
    // (down to the first prefix, or to the first one not yet published, which
    // the next round reloads), so a walk over t tiles waits on about
    // t / kLookBatch round trips instead of t.
    uint32_t before = 0;
    if (tile) {
        constexpr uint32_t kLookBatch = 8;
        uint32_t pred = tile - 1, spins = 0;
        for (;;) {
            uint32_t w[kLookBatch];
#pragma unroll
            for (uint32_t q = 0; q < kLookBatch; ++q)
                w[q] = q <= pred ? __hip_atomic_load(status + (size_t)(pred - q) * 256 + d, __ATOMIC_RELAXED,
                                                     __HIP_MEMORY_SCOPE_AGENT)
                                 : 0u;
            uint32_t used = 0;
            bool done = false;
#pragma unroll
            for (uint32_t q = 0; q < kLookBatch; ++q) {
                if (done || used < q || q > pred || (w[q] & ~kLookCount) == 0) continue;
                before += w[q] & kLookCount;
                used = q + 1;
                done = (w[q] & ~kLookCount) == kLookPrefix || pred == q;
            }
            if (done) break;
            if (used == 0 && ++spins > (1u << 26)) break;  // unreachable by construction; never hang
            pred -= used;
        }
        __hip_atomic_store(word, kLookPrefix | (before + own), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
    __syncthreads();
    uint32_t waveOffset = 0;
    for (int w = 0; w < wave; ++w) waveOffset += waveSums[w];
    running[d] = waveOffset + x - c + before;
    __syncthreads();
    const uint64_t below = lane ? (~0ull >> (64 - lane)) : 0ull;
#pragma unroll
    for (int r = 0; r < kSortItems; ++r) {
        const uint32_t i = base + r * kSortThreads + threadIdx.x;
        const bool valid = i < n;
        const uint64_t key = kr[r];
        const uint32_t val = vr[r];
        uint32_t dg = valid ? digitOf(key, shift) : 0;
        // Lanes holding the same digit: intersect eight bit-plane ballots.
        uint64_t match = __ballot(valid);
        for (int b = 0; b < 8; ++b) {
            uint64_t ones = __ballot(valid && ((dg >> b) & 1u));
            match &= ((dg >> b) & 1u) ? ones : ~ones;
        }
        const uint32_t rankInWave = (uint32_t)__popcll(match & below);
        if (valid && rankInWave == 0) waveCount[wave][dg] = (uint32_t)__popcll(match);
        __syncthreads();
        if (valid) {
            uint32_t pos = running[dg] + rankInWave;
            for (int w = 0; w < wave; ++w) pos += waveCount[w][dg];
            keysOut[pos] = key;
            valsOut[pos] = val;
        }
        __syncthreads();
        uint32_t add = 0;
        for (int w = 0; w < kSortWaves; ++w) { add += waveCount[w][d]; waveCount[w][d] = 0; }
        running[d] += add;
        __syncthreads();
    }
}

__global__ void kSortClear(uint32_t* p, uint32_t n) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = 0u;
}

// ----------------------------------------------------------------- scan --
constexpr int kScanThreads = 256;
constexpr int kScanItems = 8;
constexpr int kScanTile = kScanThreads * kScanItems;

__device__ inline uint32_t blockExclusiveScan(uint32_t v, uint32_t* smem, uint32_t& total) {
    // Hillis-Steele over the block in LDS (256 entries).
    smem[threadIdx.x] = v;
    __syncthreads();
    for (int off = 1; off < kScanThreads; off <<= 1) {
        uint32_t t = threadIdx.x >= (uint32_t)off ? smem[threadIdx.x - off] : 0u;
        __syncthreads();
        smem[threadIdx.x] += t;
        __syncthreads();
    }
    total = smem[kScanThreads - 1];
    uint32_t excl = smem[threadIdx.x] - v;
    __syncthreads();
    return excl;
}

__global__ __launch_bounds__(kScanThreads) void kScanTiles(const uint32_t* __restrict__ in, uint32_t n,
                                                           uint32_t* __restrict__ out, uint32_t* __restrict__ tileSums) {
    __shared__ uint32_t smem[kScanThreads];
    const uint32_t base = blockIdx.x * kScanTile + threadIdx.x * kScanItems;
    uint32_t local[kScanItems];
    uint32_t sum = 0;
    for (int k = 0; k < kScanItems; ++k) {
        uint32_t i = base + k;
        local[k] = i < n ? in[i] : 0u;
        sum += local[k];
    }
    uint32_t total;
    uint32_t run = blockExclusiveScan(sum, smem, total);
    for (int k = 0; k < kScanItems; ++k) {
        uint32_t i = base + k;
        if (i < n) out[i] = run;
        run += local[k];
    }
    if (threadIdx.x == 0) tileSums[blockIdx.x] = total;
}

// Add-back with each tile's offset summed from the raw tile sums by the
// tile's own workgroup (tile sums of earlier tiles, reduced in LDS), so the
// scan of the tile sums needs no launch of its own; the last tile writes the
// total.
__global__ __launch_bounds__(kScanThreads) void kScanAdd(uint32_t* __restrict__ out, uint32_t n,
                                                         const uint32_t* __restrict__ tileSums, uint32_t tiles,
                                                         uint32_t* __restrict__ total) {
    __shared__ uint32_t red[kScanThreads];
    uint32_t acc = 0;
    for (uint32_t t = threadIdx.x; t < blockIdx.x; t += kScanThreads) acc += tileSums[t];
    red[threadIdx.x] = acc;
    __syncthreads();
    for (int off = kScanThreads / 2; off > 0; off >>= 1) {
        if (threadIdx.x < (uint32_t)off) red[threadIdx.x] += red[threadIdx.x + off];
        __syncthreads();
    }
    const uint32_t add = red[0];
    const uint32_t base = blockIdx.x * kScanTile + threadIdx.x * kScanItems;
    for (int k = 0; k < kScanItems; ++k) {
        uint32_t i = base + k;
        if (i < n) out[i] += add;
    }
    if (total && threadIdx.x == 0 && blockIdx.x == tiles - 1) *total = add + tileSums[tiles - 1];
}

// Single-workgroup exclusive scan of up to kScanOneMax counts in one launch
// (the three-kernel scan above pays two extra launches, which dominate for
// small arrays such as a sort pass's histogram table).  Each thread scans 16
// consecutive counts in registers; wavefront and workgroup offsets come from
// lane shuffles and 16 LDS words.  `out` may alias `in`: every thread reads
// its own counts before writing them.
constexpr int kScanOneThreads = 1024;
constexpr int kScanOneItems = 16;
constexpr uint32_t kScanOneMax = kScanOneThreads * kScanOneItems;

__global__ __launch_bounds__(kScanOneThreads) void kScanOne(const uint32_t* in, uint32_t n, uint32_t* out,
                                                            uint32_t* __restrict__ total) {
    __shared__ uint32_t waveSums[kScanOneThreads / 64];
    const uint32_t t = threadIdx.x, lane = t & 63, wave = t >> 6;
    const uint32_t base = t * kScanOneItems;
    uint32_t local[kScanOneItems];
    uint32_t sum = 0;
#pragma unroll
    for (int k = 0; k < kScanOneItems; ++k) {
        local[k] = base + k < n ? in[base + k] : 0u;
        sum += local[k];
    }
    uint32_t x = sum;  // inclusive scan over the wavefront
#pragma unroll
    for (int off = 1; off < 64; off <<= 1) {
        const uint32_t y = __shfl_up(x, off, 64);
        if (lane >= (uint32_t)off) x += y;
    }
    if (lane == 63) waveSums[wave] = x;
    __syncthreads();
    if (wave == 0) {
        uint32_t w = lane < kScanOneThreads / 64 ? waveSums[lane] : 0u;
#pragma unroll
        for (int off = 1; off < kScanOneThreads / 64; off <<= 1) {
            const uint32_t y = __shfl_up(w, off, 64);
            if (lane >= (uint32_t)off) w += y;
        }
        if (lane < kScanOneThreads / 64) waveSums[lane] = w;
    }
    __syncthreads();
    const uint32_t waveOffset = wave ? waveSums[wave - 1] : 0u;
    uint32_t run = waveOffset + x - sum;
#pragma unroll
    for (int k = 0; k < kScanOneItems; ++k) {
        if (base + k < n) out[base + k] = run;
        run += local[k];
    }
    if (t == kScanOneThreads - 1 && total) *total = waveOffset + x;
}

// ------------------------------------------------ exact std::sort order --
// libstdc++'s std::sort (introsort: median-of-three pivot moved to the front,
// unguarded Hoare partition of the rest, 16-element threshold, heap sort
// after 2*floor(log2 n) levels, final insertion sort) leaves equal keys in an
// order that depends on the whole input.  The pending ranges of the
// introsort touch disjoint elements, so they are processed level by level,
// one workgroup per range, and each Hoare partition is computed from ranks:
// with L_k the k-th position (from the left) whose key is not below the
// pivot and R_k the k-th position (from the right) whose key is not above
// it, the partition swaps (L_k, R_k) exactly for the k with L_k < R_k (a
// prefix, since L rises and R falls) and returns min(L_s, R_{s-1}), s the
// number of swaps.  Between two swaps each scanner only reads elements no
// earlier swap has touched, which is what makes the ranks of the input
// enough.  After the levels, every range of at most 16 elements lies between
// smaller and larger ranges, so the final insertion sort moves elements only
// within their range: a stable insertion sort per range finishes the order.
constexpr int kIntroThreads = 256;
constexpr int kIntroItems = 4;
constexpr uint32_t kIntroTile = kIntroThreads * kIntroItems;
constexpr uint32_t kIntroSorted = 0x80000000u;  // leaf already sorted (heap sort)

struct IntroSeg {
    uint32_t first, last, depth, pad;
};

// Exclusive scan of one packed u32 per thread over the workgroup.
__device__ inline uint32_t introScan(uint32_t v, uint32_t* waveSums, uint32_t& total) {
    const uint32_t lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    uint32_t x = v;
#pragma unroll
    for (int off = 1; off < 64; off <<= 1) {
        const uint32_t y = __shfl_up(x, off, 64);
        if (lane >= (uint32_t)off) x += y;
    }
    if (lane == 63) waveSums[wave] = x;
    __syncthreads();
    uint32_t before = 0, all = 0;
#pragma unroll
    for (uint32_t w = 0; w < kIntroThreads / 64; ++w) {
        const uint32_t t = waveSums[w];
        before += w < wave ? t : 0u;
        all += t;
    }
    __syncthreads();
    total = all;
    return before + x - v;
}

__device__ inline void introEmit(uint32_t a, uint32_t b, uint32_t depth, IntroSeg* next, uint32_t* nextCount,
                                 uint2* leaves, uint32_t* leafCount) {
    if (b - a > vxsort::kThreshold) {
        next[atomicAdd(nextCount, 1u)] = IntroSeg{a, b, depth, 0u};
    } else if (b - a > 1) {
        leaves[atomicAdd(leafCount, 1u)] = make_uint2(a, b);
    }
}

__global__ void kIntroClear(uint32_t* counts, uint32_t n) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) counts[i] = 0u;
}

// Level 0: one range per segment, with std::sort's depth limit for its size.
// Segment offsets that do not rise from 0 to n without falling (a device
// array the host has not seen) set *bad; kIntroInit then emits nothing.
__global__ void kIntroCheck(const uint32_t* __restrict__ segStart, uint32_t segments, uint32_t n,
                            uint32_t* __restrict__ bad) {
    const uint32_t w = blockIdx.x * blockDim.x + threadIdx.x;
    if (w >= segments) return;
    const bool ok = segStart[w] <= segStart[w + 1] && (w > 0 || segStart[0] == 0) &&
                    (w + 1 < segments || segStart[segments] == n);
    if (!ok) atomicOr(bad, 1u);
}

__global__ void kIntroInit(const uint32_t* __restrict__ segStart, uint32_t segments, uint32_t n,
                           IntroSeg* __restrict__ level0, uint32_t* __restrict__ count0,
                           uint2* __restrict__ leaves, uint32_t* __restrict__ leafCount,
                           const uint32_t* __restrict__ bad) {
    const uint32_t w = blockIdx.x * blockDim.x + threadIdx.x;
    if (w >= segments || (bad && *bad)) return;
    const uint32_t f = segStart ? segStart[w] : 0u, l = segStart ? segStart[w + 1] : n;
    if (l <= f) return;
    introEmit(f, l, 2u * vxsort::floorLog2(l - f), level0, count0, leaves, leafCount);
}

__global__ __launch_bounds__(kIntroThreads) void kIntroLevel(uint64_t* __restrict__ K, uint32_t* __restrict__ V,
                                                             const IntroSeg* __restrict__ segs,
                                                             const uint32_t* __restrict__ segCount,
                                                             IntroSeg* __restrict__ next, uint32_t* __restrict__ nextCount,
                                                             uint32_t* __restrict__ posL, uint32_t* __restrict__ posR,
                                                             uint2* __restrict__ leaves, uint32_t* __restrict__ leafCount) {
    __shared__ uint32_t waveSums[kIntroThreads / 64];
    __shared__ uint64_t pivotS;
    __shared__ uint32_t swapsS;
    const vxsort::Seq seq{K, V};
    const uint32_t count = *segCount;
    for (uint32_t si = blockIdx.x; si < count; si += gridDim.x) {
        const IntroSeg g = segs[si];
        const uint32_t f = g.first, l = g.last;
        if (g.depth == 0) {  // __partial_sort(first, last, last)
            if (threadIdx.x == 0) {
                vxsort::heapSort(seq, f, l);
                leaves[atomicAdd(leafCount, 1u)] = make_uint2(f, l | kIntroSorted);
            }
            __syncthreads();
            continue;
        }
        if (threadIdx.x == 0) {  // __move_median_to_first(first, first + 1, mid, last - 1)
            vxsort::medianToFirst(seq, f, f + 1, f + (l - f) / 2, l - 1);
            pivotS = K[f];
        }
        __syncthreads();
        const uint64_t P = pivotS;
        // Stops of both scanners over [f + 1, l), in increasing position.
        uint32_t cl = 0, cr = 0;
        for (uint32_t base = f + 1; base < l; base += kIntroTile) {
            const uint32_t p0 = base + threadIdx.x * kIntroItems;
            uint32_t fl = 0, fr = 0;
#pragma unroll
            for (int j = 0; j < kIntroItems; ++j) {
                const uint32_t p = p0 + j;
                if (p < l) {
                    const uint64_t k = K[p];
                    fl |= (k < P ? 0u : 1u) << j;
                    fr |= (P < k ? 0u : 1u) << j;
                }
            }
            uint32_t total;
            const uint32_t ex = introScan((uint32_t)__popc(fl) | ((uint32_t)__popc(fr) << 16), waveSums, total);
            uint32_t rl = cl + (ex & 0xffffu), rr = cr + (ex >> 16);
#pragma unroll
            for (int j = 0; j < kIntroItems; ++j) {
                if ((fl >> j) & 1u) posL[f + rl++] = p0 + j;
                if ((fr >> j) & 1u) posR[f + rr++] = p0 + j;
            }
            cl += total & 0xffffu;
            cr += total >> 16;
        }
        __syncthreads();
        // s = #{k : L_k < R_k}, R_k = posR[f + cr - 1 - k]: a prefix of k.
        if (threadIdx.x == 0) {
            uint32_t lo = 0, hi = cl < cr ? cl : cr;
            while (lo < hi) {
                const uint32_t mid = (lo + hi) / 2;
                if (posL[f + mid] < posR[f + cr - 1 - mid]) lo = mid + 1;
                else hi = mid;
            }
            swapsS = lo;
        }
        __syncthreads();
        const uint32_t swaps = swapsS;
        for (uint32_t k = threadIdx.x; k < swaps; k += kIntroThreads)
            seq.swap(posL[f + k], posR[f + cr - 1 - k]);
        __syncthreads();
        if (threadIdx.x == 0) {
            uint32_t cut = swaps ? posR[f + cr - swaps] : l;
            if (swaps < cl && posL[f + swaps] < cut) cut = posL[f + swaps];
            introEmit(f, cut, g.depth - 1, next, nextCount, leaves, leafCount);
            introEmit(cut, l, g.depth - 1, next, nextCount, leaves, leafCount);
        }
        __syncthreads();
    }
}

// The final insertion sort, range by range (stable insertion sort).
__global__ void kIntroLeaves(uint64_t* __restrict__ K, uint32_t* __restrict__ V, const uint2* __restrict__ leaves,
                             const uint32_t* __restrict__ leafCount) {
    const vxsort::Seq seq{K, V};
    const uint32_t count = *leafCount;
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < count; i += gridDim.x * blockDim.x) {
        const uint2 r = leaves[i];
        if (r.y & kIntroSorted) continue;
        vxsort::insertionSort(seq, r.x, r.y);
    }
}

}  // namespace vxk


namespace vxhost {

using namespace vxk;

// Working memory of a sort: per pass the 256 digit counts, the tile counter
// and the tiles' look-back words.
size_t radixTempBytes(uint32_t n) {
    const uint32_t tiles = (n + kSortTile - 1) / kSortTile;
    return (size_t)kMaxPasses * (256 + 1 + 256 * (size_t)(tiles ? tiles : 1)) * sizeof(uint32_t) + 64;
}

// Sorts (keys, vals) in place by the low `bits` of the key.  `keysTmp`,
// `valsTmp` hold n elements; `temp` holds radixTempBytes(n) bytes.  With
// `nDev` the count is read on the device and n is the capacity: the grids
// are sized for n and the workgroups past the count return at once, so no
// host round trip is needed to learn the count.
hipError_t radixSort(uint64_t* keys, uint32_t* vals, uint64_t* keysTmp, uint32_t* valsTmp, uint32_t n, int bits,
                     void* temp, hipStream_t s, const uint32_t* nDev) {
    if ((n <= 1 && !nDev) || n == 0 || bits <= 0) return hipSuccess;
    const uint32_t tiles = (n + kSortTile - 1) / kSortTile;
    const int passes = std::min((bits + 7) / 8, kMaxPasses);
    uint32_t* digitCounts = (uint32_t*)temp;                            // passes x 256
    uint32_t* tileCounter = digitCounts + (size_t)passes * 256;         // passes
    uint32_t* status = tileCounter + passes;                            // passes x tiles x 256
    const uint32_t words = (uint32_t)((size_t)passes * (257 + 256 * (size_t)tiles));
    kSortClear<<<(words + 255) / 256, 256, 0, s>>>(digitCounts, words);
    // An odd pass count starts from a copy in the second buffer (made by the
    // counting launch), so the passes end in (keys, vals).
    const bool odd = passes & 1;
    kRadixDigitCounts<<<tiles, kSortThreads, 0, s>>>(keys, vals, n, nDev, passes, digitCounts,
                                                     odd ? keysTmp : nullptr, odd ? valsTmp : nullptr);
    uint64_t* ka = odd ? keysTmp : keys; uint32_t* va = odd ? valsTmp : vals;
    uint64_t* kb = odd ? keys : keysTmp; uint32_t* vb = odd ? vals : valsTmp;
    for (int p = 0; p < passes; ++p) {
        kRadixPass<<<tiles, kSortThreads, 0, s>>>(ka, va, n, nDev, 8 * p, digitCounts + 256 * p, tileCounter + p,
                                                  status + (size_t)p * tiles * 256, kb, vb);
        uint64_t* tk = ka; ka = kb; kb = tk;
        uint32_t* tv = va; va = vb; vb = tv;
    }
    return hipGetLastError();
}

size_t scanTempBytes(uint32_t n) { return ((size_t)(n + kScanTile - 1) / kScanTile + 1) * sizeof(uint32_t) + 64; }

// Exclusive scan of n counts into out (may alias in); *total receives the sum.
hipError_t exclusiveScan(const uint32_t* in, uint32_t* out, uint32_t n, uint32_t* total, void* temp, hipStream_t s) {
    if (n == 0) {
        (void)hipMemsetAsync(total, 0, sizeof(uint32_t), s);
        return hipGetLastError();
    }
    if (n <= kScanOneMax) {
        kScanOne<<<1, kScanOneThreads, 0, s>>>(in, n, out, total);
        return hipGetLastError();
    }
    const uint32_t tiles = (n + kScanTile - 1) / kScanTile;
    uint32_t* sums = (uint32_t*)temp;
    kScanTiles<<<tiles, kScanThreads, 0, s>>>(in, n, out, sums);
    kScanAdd<<<tiles, kScanThreads, 0, s>>>(out, n, sums, tiles, total);
    return hipGetLastError();
}

namespace {
struct IntroLayout {
    uint32_t maxSegs, levels;
    size_t posL, posR, segA, segB, counts, leaves, bytes;
};
IntroLayout introLayout(uint32_t n, uint32_t segments) {
    IntroLayout L{};
    L.maxSegs = std::max<uint32_t>(segments, n / (vxsort::kThreshold + 1)) + 2;
    L.levels = 2 * vxsort::floorLog2(n ? n : 1) + 2;
    size_t at = 0;
    auto take = [&](size_t bytes) { const size_t p = at; at += (bytes + 255) & ~(size_t)255; return p; };
    L.posL = take((size_t)n * 4);
    L.posR = take((size_t)n * 4);
    L.segA = take((size_t)L.maxSegs * sizeof(IntroSeg));
    L.segB = take((size_t)L.maxSegs * sizeof(IntroSeg));
    L.counts = take((size_t)(L.levels + 2) * 4);
    L.leaves = take((size_t)(n / 2 + segments + 2) * sizeof(uint2));
    L.bytes = at;
    return L;
}
}  // namespace

size_t stdSortTempBytes(uint32_t n, uint32_t segments) { return introLayout(n, segments).bytes; }

// Orders (keys, vals) exactly as std::sort would, segment by segment.
hipError_t stdSort(uint64_t* keys, uint32_t* vals, uint32_t n, const uint32_t* segStart, uint32_t segments,
                   void* temp, hipStream_t s, uint32_t* bad) {
    if (n <= 1 || segments == 0) return hipSuccess;
    const IntroLayout L = introLayout(n, segments);
    char* t = (char*)temp;
    uint32_t* posL = (uint32_t*)(t + L.posL);
    uint32_t* posR = (uint32_t*)(t + L.posR);
    IntroSeg* seg[2] = {(IntroSeg*)(t + L.segA), (IntroSeg*)(t + L.segB)};
    uint32_t* counts = (uint32_t*)(t + L.counts);  // per level, then the leaf count
    uint2* leaves = (uint2*)(t + L.leaves);
    uint32_t* leafCount = counts + L.levels + 1;
    const uint32_t nc = L.levels + 2;
    kIntroClear<<<(nc + 255) / 256, 256, 0, s>>>(counts, nc);
    if (bad && segStart) kIntroCheck<<<(segments + 255) / 256, 256, 0, s>>>(segStart, segments, n, bad);
    kIntroInit<<<(segments + 255) / 256, 256, 0, s>>>(segStart, segments, n, seg[0], counts, leaves, leafCount,
                                                      segStart ? bad : nullptr);
    for (uint32_t level = 0; level + 1 < L.levels; ++level) {
        const uint64_t wide = (uint64_t)segments << std::min<uint32_t>(level, 24);
        const uint32_t grid = (uint32_t)std::min<uint64_t>(std::min<uint64_t>(wide, L.maxSegs), 4096);
        kIntroLevel<<<grid, kIntroThreads, 0, s>>>(keys, vals, seg[level & 1], counts + level, seg[(level + 1) & 1],
                                                   counts + level + 1, posL, posR, leaves, leafCount);
    }
    const uint32_t leafGrid = std::min<uint32_t>((n / 2 + segments + 257) / 256, 2048);
    kIntroLeaves<<<leafGrid, 256, 0, s>>>(keys, vals, leaves, leafCount);
    return hipGetLastError();
}

}  // namespace vxhost

