// Internal declarations shared by the HIP translation units of libvxhip.
#pragma once
#include <hip/hip_runtime.h>
#include <cstddef>
#include <cstdint>

namespace vxhost {

size_t radixTempBytes(uint32_t n);
// With nDev, the count is read on the device and n is only the capacity.
hipError_t radixSort(uint64_t* keys, uint32_t* vals, uint64_t* keysTmp, uint32_t* valsTmp, uint32_t n, int bits,
                     void* temp, hipStream_t s, const uint32_t* nDev = nullptr);
size_t scanTempBytes(uint32_t n);
hipError_t exclusiveScan(const uint32_t* in, uint32_t* out, uint32_t n, uint32_t* total, void* temp, hipStream_t s);

// libstdc++'s std::sort permutation of (key, value) pairs, reproduced exactly
// and in parallel (vx_sort.hip).  The input is cut into independent segments
// [segStart[w], segStart[w + 1]) (one per world; segStart == nullptr: one
// segment [0, n)), each sorted as its own std::sort call would sort it.
// With `bad` (a zeroed device word), segment offsets that do not rise from
// 0 to n are detected on the device: *bad is set and nothing is sorted.
size_t stdSortTempBytes(uint32_t n, uint32_t segments);
hipError_t stdSort(uint64_t* keys, uint32_t* vals, uint32_t n, const uint32_t* segStart, uint32_t segments,
                   void* temp, hipStream_t s, uint32_t* bad = nullptr);

inline int bitsFor(uint64_t maxValue) {  // bits needed to represent 0..maxValue
    int b = 0;
    while (b < 64 && (maxValue >> b) != 0) ++b;
    return b;
}

}  // namespace vxhost
