// libstdc++'s std::sort (bits/stl_algo.h, GCC 11) restated over (key, value)
// pairs compared by key alone: introsort with median-of-three pivots moved to
// the front, Hoare partitioning of the rest, a 16-element threshold, heap sort
// once 2*floor(log2 n) levels are spent, then insertion sort.  std::sort is
// not stable, and Velox's World::step orders its warm-start cache with it
// (world.cpp:4518-4521); which of a pair's cached contacts comes first there
// decides the contact an inert pair's Persist event reports and the one
// contact kept for it.  This restatement reproduces that order exactly.
//
// Plain C++ (one thread): compiled into the device library and, for its test
// against std::sort itself, by the host compiler.
#pragma once
#include <cstdint>

#if defined(__HIPCC__)
#define VXSS __host__ __device__ inline
#else
#define VXSS inline
#endif

namespace vxsort {

struct Seq {
    uint64_t* k;
    uint32_t* v;
    VXSS bool less(uint32_t i, uint32_t j) const { return k[i] < k[j]; }
    VXSS void swap(uint32_t i, uint32_t j) const {
        const uint64_t tk = k[i]; k[i] = k[j]; k[j] = tk;
        const uint32_t tv = v[i]; v[i] = v[j]; v[j] = tv;
    }
    VXSS void move(uint32_t to, uint32_t from) const { k[to] = k[from]; v[to] = v[from]; }
};

VXSS uint32_t floorLog2(uint32_t n) {
    uint32_t r = 0;
    while (n > 1) { n >>= 1; ++r; }
    return r;
}

// __move_median_to_first
VXSS void medianToFirst(const Seq& s, uint32_t result, uint32_t a, uint32_t b, uint32_t c) {
    if (s.less(a, b)) {
        if (s.less(b, c)) s.swap(result, b);
        else if (s.less(a, c)) s.swap(result, c);
        else s.swap(result, a);
    } else if (s.less(a, c)) s.swap(result, a);
    else if (s.less(b, c)) s.swap(result, c);
    else s.swap(result, b);
}

// __unguarded_partition(first, last, pivot)
VXSS uint32_t partition(const Seq& s, uint32_t first, uint32_t last, uint32_t pivot) {
    for (;;) {
        while (s.less(first, pivot)) ++first;
        --last;
        while (s.less(pivot, last)) --last;
        if (!(first < last)) return first;
        s.swap(first, last);
        ++first;
    }
}

// __adjust_heap + __push_heap over [base, base + len) with the value (vk, vv).
VXSS void adjustHeap(const Seq& s, uint32_t base, uint32_t hole, uint32_t len, uint64_t vk, uint32_t vv) {
    const uint32_t top = hole;
    uint32_t child = hole;
    while (len >= 1 && child < (len - 1) / 2) {
        child = 2 * (child + 1);
        if (s.less(base + child, base + (child - 1))) --child;
        s.move(base + hole, base + child);
        hole = child;
    }
    if ((len & 1) == 0 && len >= 2 && child == (len - 2) / 2) {
        child = 2 * (child + 1);
        s.move(base + hole, base + (child - 1));
        hole = child - 1;
    }
    uint32_t parent = hole ? (hole - 1) / 2 : 0;
    while (hole > top && s.k[base + parent] < vk) {
        s.move(base + hole, base + parent);
        hole = parent;
        parent = hole ? (hole - 1) / 2 : 0;
    }
    s.k[base + hole] = vk;
    s.v[base + hole] = vv;
}

// __partial_sort(first, last, last): __make_heap then __sort_heap.
VXSS void heapSort(const Seq& s, uint32_t first, uint32_t last) {
    const uint32_t len = last - first;
    if (len >= 2) {
        uint32_t parent = (len - 2) / 2;
        for (;;) {
            adjustHeap(s, first, parent, len, s.k[first + parent], s.v[first + parent]);
            if (parent == 0) break;
            --parent;
        }
    }
    while (last - first > 1) {
        --last;
        const uint64_t vk = s.k[last];
        const uint32_t vv = s.v[last];
        s.move(last, first);
        adjustHeap(s, first, 0, last - first, vk, vv);
    }
}

// __unguarded_linear_insert
VXSS void linearInsert(const Seq& s, uint32_t last) {
    const uint64_t vk = s.k[last];
    const uint32_t vv = s.v[last];
    uint32_t next = last - 1;
    while (vk < s.k[next]) {
        s.move(last, next);
        last = next;
        --next;
    }
    s.k[last] = vk;
    s.v[last] = vv;
}

// __insertion_sort
VXSS void insertionSort(const Seq& s, uint32_t first, uint32_t last) {
    if (first == last) return;
    for (uint32_t i = first + 1; i != last; ++i) {
        if (s.less(i, first)) {
            const uint64_t vk = s.k[i];
            const uint32_t vv = s.v[i];
            for (uint32_t j = i; j > first; --j) s.move(j, j - 1);
            s.k[first] = vk;
            s.v[first] = vv;
        } else {
            linearInsert(s, i);
        }
    }
}

constexpr uint32_t kThreshold = 16;  // _S_threshold

// std::sort(first, first + n).  The introsort's recursion on the right part
// and loop on the left touch disjoint ranges, so the pending ranges are kept
// on a small explicit stack (larger part pushed, smaller continued: at most
// log2(n) + 1 entries).
VXSS void sort(const Seq& s, uint32_t n) {
    if (n < 2) return;
    struct Range { uint32_t first, last, depth; };
    Range stack[40];
    int top = 0;
    stack[top++] = {0u, n, 2u * floorLog2(n)};
    while (top > 0) {
        Range r = stack[--top];
        while (r.last - r.first > kThreshold) {
            if (r.depth == 0) { heapSort(s, r.first, r.last); break; }
            --r.depth;
            const uint32_t mid = r.first + (r.last - r.first) / 2;
            medianToFirst(s, r.first, r.first + 1, mid, r.last - 1);
            const uint32_t cut = partition(s, r.first + 1, r.last, r.first);
            const Range right{cut, r.last, r.depth}, left{r.first, cut, r.depth};
            if (right.last - right.first > left.last - left.first) { stack[top++] = right; r = left; }
            else { stack[top++] = left; r = right; }
        }
    }
    if (n > kThreshold) {
        insertionSort(s, 0, kThreshold);
        for (uint32_t i = kThreshold; i != n; ++i) linearInsert(s, i);
    } else {
        insertionSort(s, 0, n);
    }
}

}  // namespace vxsort
