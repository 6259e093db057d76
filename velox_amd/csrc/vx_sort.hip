// Hand-written device sort and scan primitives for the HIP device step.
//
//  * LSD radix sort of (u64 key, u32 value) pairs over the low `bits` of the
//    key, 8-bit digits.  Each pass: per-tile digit histograms (LDS), one scan
//    over the (digit-major, tile-minor) histogram table, then a stable scatter
//    in which every lane ranks itself among equal digits of its wavefront with
//    eight ballots (a 64-lane match mask) and waves are chained through LDS.
//    Stability per pass is what makes the multi-pass LSD sort correct.
//  * Exclusive scan of u32 counts: one workgroup up to 16384 counts, else
//    tile sums, scan of tile sums, add-back.
#include "vx_internal.hpp"

namespace vxk {

constexpr int kSortThreads = 256;          // 4 wavefronts
constexpr int kSortWaves = kSortThreads / 64;
constexpr int kSortItems = 16;             // rounds per tile
constexpr int kSortTile = kSortThreads * kSortItems;

__device__ inline uint32_t digitOf(uint64_t key, int shift) { return (uint32_t)(key >> shift) & 0xffu; }

__global__ __launch_bounds__(kSortThreads) void kRadixHistogram(const uint64_t* __restrict__ keys, uint32_t n,
                                                                int shift, uint32_t* __restrict__ hist,
                                                                uint32_t tiles) {
    __shared__ uint32_t counts[256];
    const uint32_t tile = blockIdx.x;
    counts[threadIdx.x] = 0;
    __syncthreads();
    const uint32_t base = tile * kSortTile;
    for (int r = 0; r < kSortItems; ++r) {
        uint32_t i = base + r * kSortThreads + threadIdx.x;
        if (i < n) atomicAdd(&counts[digitOf(keys[i], shift)], 1u);
    }
    __syncthreads();
    hist[threadIdx.x * tiles + tile] = counts[threadIdx.x];
}

__global__ __launch_bounds__(kSortThreads) void kRadixScatter(const uint64_t* __restrict__ keysIn,
                                                              const uint32_t* __restrict__ valsIn, uint32_t n,
                                                              int shift, const uint32_t* __restrict__ offsets,
                                                              uint32_t tiles, uint64_t* __restrict__ keysOut,
                                                              uint32_t* __restrict__ valsOut) {
    __shared__ uint32_t running[256];
    __shared__ uint32_t waveCount[kSortWaves][256];
    const uint32_t tile = blockIdx.x;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    running[threadIdx.x] = offsets[threadIdx.x * tiles + tile];
    for (int w = 0; w < kSortWaves; ++w) waveCount[w][threadIdx.x] = 0;
    __syncthreads();
    const uint64_t below = lane ? (~0ull >> (64 - lane)) : 0ull;
    const uint32_t base = tile * kSortTile;
    for (int r = 0; r < kSortItems; ++r) {
        const uint32_t i = base + r * kSortThreads + threadIdx.x;
        const bool valid = i < n;
        uint64_t key = valid ? keysIn[i] : 0;
        uint32_t val = valid ? valsIn[i] : 0;
        uint32_t d = valid ? digitOf(key, shift) : 0;
        // Lanes holding the same digit: intersect eight bit-plane ballots.
        uint64_t match = __ballot(valid);
        for (int b = 0; b < 8; ++b) {
            uint64_t ones = __ballot(valid && ((d >> b) & 1u));
            match &= ((d >> b) & 1u) ? ones : ~ones;
        }
        const uint32_t rankInWave = (uint32_t)__popcll(match & below);
        if (valid && rankInWave == 0) waveCount[wave][d] = (uint32_t)__popcll(match);
        __syncthreads();
        if (valid) {
            uint32_t pos = running[d] + rankInWave;
            for (int w = 0; w < wave; ++w) pos += waveCount[w][d];
            keysOut[pos] = key;
            valsOut[pos] = val;
        }
        __syncthreads();
        uint32_t add = 0;
        for (int w = 0; w < kSortWaves; ++w) { add += waveCount[w][threadIdx.x]; waveCount[w][threadIdx.x] = 0; }
        running[threadIdx.x] += add;
        __syncthreads();
    }
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

// Single-block scan of the tile sums (any count) and total.
__global__ __launch_bounds__(kScanThreads) void kScanSums(uint32_t* __restrict__ sums, uint32_t count,
                                                          uint32_t* __restrict__ totalOut) {
    __shared__ uint32_t smem[kScanThreads];
    uint32_t carry = 0;
    for (uint32_t start = 0; start < count; start += kScanThreads) {
        uint32_t i = start + threadIdx.x;
        uint32_t v = i < count ? sums[i] : 0u;
        uint32_t total;
        uint32_t ex = blockExclusiveScan(v, smem, total);
        if (i < count) sums[i] = carry + ex;
        carry += total;
    }
    if (threadIdx.x == 0 && totalOut) *totalOut = carry;
}

__global__ __launch_bounds__(kScanThreads) void kScanAdd(uint32_t* __restrict__ out, uint32_t n,
                                                         const uint32_t* __restrict__ tileSums) {
    const uint32_t base = blockIdx.x * kScanTile + threadIdx.x * kScanItems;
    const uint32_t add = tileSums[blockIdx.x];
    for (int k = 0; k < kScanItems; ++k) {
        uint32_t i = base + k;
        if (i < n) out[i] += add;
    }
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

}  // namespace vxk

namespace vxhost {

using namespace vxk;

size_t radixTempBytes(uint32_t n) {
    uint32_t tiles = (n + kSortTile - 1) / kSortTile;
    return (size_t)256 * (tiles ? tiles : 1) * sizeof(uint32_t) * 2 + 64;
}

// Sorts (keys, vals) in place by the low `bits` of the key.  `keysTmp`,
// `valsTmp` hold n elements; `temp` holds radixTempBytes(n) bytes.
hipError_t radixSort(uint64_t* keys, uint32_t* vals, uint64_t* keysTmp, uint32_t* valsTmp, uint32_t n, int bits,
                     void* temp, hipStream_t s) {
    if (n <= 1 || bits <= 0) return hipSuccess;
    const uint32_t tiles = (n + kSortTile - 1) / kSortTile;
    uint32_t* hist = (uint32_t*)temp;
    uint32_t* sums = hist + (size_t)256 * tiles;
    uint64_t* ka = keys; uint32_t* va = vals;
    uint64_t* kb = keysTmp; uint32_t* vb = valsTmp;
    const uint32_t histN = 256 * tiles;
    const uint32_t scanTiles = (histN + kScanTile - 1) / kScanTile;
    int passes = (bits + 7) / 8;
    for (int p = 0; p < passes; ++p) {
        const int shift = 8 * p;
        kRadixHistogram<<<tiles, kSortThreads, 0, s>>>(ka, n, shift, hist, tiles);
        if (histN <= kScanOneMax) {
            kScanOne<<<1, kScanOneThreads, 0, s>>>(hist, histN, hist, nullptr);
        } else {
            kScanTiles<<<scanTiles, kScanThreads, 0, s>>>(hist, histN, hist, sums);
            kScanSums<<<1, kScanThreads, 0, s>>>(sums, scanTiles, nullptr);
            kScanAdd<<<scanTiles, kScanThreads, 0, s>>>(hist, histN, sums);
        }
        kRadixScatter<<<tiles, kSortThreads, 0, s>>>(ka, va, n, shift, hist, tiles, kb, vb);
        uint64_t* tk = ka; ka = kb; kb = tk;
        uint32_t* tv = va; va = vb; vb = tv;
    }
    if (ka != keys) {
        (void)hipMemcpyAsync(keys, ka, (size_t)n * sizeof(uint64_t), hipMemcpyDeviceToDevice, s);
        (void)hipMemcpyAsync(vals, va, (size_t)n * sizeof(uint32_t), hipMemcpyDeviceToDevice, s);
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
    kScanSums<<<1, kScanThreads, 0, s>>>(sums, tiles, total);
    kScanAdd<<<tiles, kScanThreads, 0, s>>>(out, n, sums);
    return hipGetLastError();
}

}  // namespace vxhost
