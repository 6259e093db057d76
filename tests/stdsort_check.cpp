// Checks the two libstdc++ std::sort restatements of velox_amd/csrc against
// std::sort and std::partial_sort of the host's libstdc++ on seeded random
// (key, index) sequences with many equal keys: the index order left among
// equal keys must be identical.
//
//  * vxsort::sort / vxsort::heapSort (vx_stdsort.hpp): the one-thread form.
//  * The level-by-level form of vx_sort.hip (kIntroInit / kIntroLevel /
//    kIntroLeaves), restated here sequentially with the same arithmetic: the
//    pending ranges of each level, each Hoare partition computed from the
//    ranks of the two scanners' stops (swap (L_k, R_k) while L_k < R_k, cut
//    at min(L_s, R_{s-1})), and a stable insertion sort per final range.  It
//    runs over several independent segments at once, each of which must come
//    out as std::sort of that segment alone.
//
// Prints the case and mismatch counts; exits non-zero on a mismatch or when
// the ranks of a partition break the prefix property the kernel relies on.
#include "../velox_amd/csrc/vx_stdsort.hpp"
#include <algorithm>
#include <cstdio>
#include <cstdlib>
#include <random>
#include <vector>

namespace {
struct E { uint64_t k; uint32_t v; };
bool byKey(const E& x, const E& y) { return x.k < y.k; }

long check(std::vector<E> a, bool heap) {
    const uint32_t n = (uint32_t)a.size();
    std::vector<uint64_t> K(n);
    std::vector<uint32_t> V(n);
    for (uint32_t i = 0; i < n; ++i) { K[i] = a[i].k; V[i] = a[i].v; }
    if (heap) {
        std::partial_sort(a.begin(), a.end(), a.end(), byKey);
        vxsort::heapSort({K.data(), V.data()}, 0, n);
    } else {
        std::sort(a.begin(), a.end(), byKey);
        vxsort::sort({K.data(), V.data()}, n);
    }
    for (uint32_t i = 0; i < n; ++i)
        if (K[i] != a[i].k || V[i] != a[i].v) return 1;
    return 0;
}

struct Seg { uint32_t first, last, depth; };

// The level-by-level form over segments [starts[w], starts[w + 1]).
void levelSort(std::vector<uint64_t>& K, std::vector<uint32_t>& V, const std::vector<uint32_t>& starts) {
    const vxsort::Seq seq{K.data(), V.data()};
    std::vector<Seg> cur, next;
    std::vector<Seg> leaves;  // depth field: 1 = already sorted
    auto emit = [&](uint32_t a, uint32_t b, uint32_t depth, std::vector<Seg>& to) {
        if (b - a > vxsort::kThreshold) to.push_back({a, b, depth});
        else if (b - a > 1) leaves.push_back({a, b, 0});
    };
    for (size_t w = 0; w + 1 < starts.size(); ++w)
        if (starts[w + 1] > starts[w]) emit(starts[w], starts[w + 1], 2 * vxsort::floorLog2(starts[w + 1] - starts[w]), cur);
    while (!cur.empty()) {
        next.clear();
        for (const Seg& g : cur) {
            const uint32_t f = g.first, l = g.last;
            if (g.depth == 0) {
                vxsort::heapSort(seq, f, l);
                leaves.push_back({f, l, 1});
                continue;
            }
            vxsort::medianToFirst(seq, f, f + 1, f + (l - f) / 2, l - 1);
            const uint64_t P = K[f];
            std::vector<uint32_t> L, R;  // both in increasing position
            for (uint32_t p = f + 1; p < l; ++p) {
                if (!(K[p] < P)) L.push_back(p);
                if (!(P < K[p])) R.push_back(p);
            }
            const uint32_t cl = (uint32_t)L.size(), cr = (uint32_t)R.size();
            uint32_t s = 0;
            while (s < std::min(cl, cr) && L[s] < R[cr - 1 - s]) ++s;
            for (uint32_t k = s; k < std::min(cl, cr); ++k)
                if (L[k] < R[cr - 1 - k]) { std::fprintf(stderr, "prefix property broken\n"); std::exit(2); }
            for (uint32_t k = 0; k < s; ++k) seq.swap(L[k], R[cr - 1 - k]);
            uint32_t cut = s ? R[cr - s] : l;
            if (s < cl && L[s] < cut) cut = L[s];
            emit(f, cut, g.depth - 1, next);
            emit(cut, l, g.depth - 1, next);
        }
        cur.swap(next);
    }
    for (const Seg& r : leaves)
        if (!r.depth) vxsort::insertionSort(seq, r.first, r.last);
}

long checkLevels(const std::vector<E>& a, const std::vector<uint32_t>& starts) {
    std::vector<E> want = a;
    for (size_t w = 0; w + 1 < starts.size(); ++w)
        std::sort(want.begin() + starts[w], want.begin() + starts[w + 1], byKey);
    std::vector<uint64_t> K(a.size());
    std::vector<uint32_t> V(a.size());
    for (size_t i = 0; i < a.size(); ++i) { K[i] = a[i].k; V[i] = a[i].v; }
    levelSort(K, V, starts);
    for (size_t i = 0; i < a.size(); ++i)
        if (K[i] != want[i].k || V[i] != want[i].v) return 1;
    return 0;
}

std::vector<E> randomCase(std::mt19937& rng, uint32_t n) {
    std::vector<E> a(n);
    const uint32_t kinds = 1 + rng() % (n + 1);
    for (uint32_t i = 0; i < n; ++i) a[i].k = rng() % kinds;
    switch (rng() % 4) {
    case 1: std::stable_sort(a.begin(), a.end(), byKey); break;
    case 2: {  // two sorted runs: the warm-start cache's shape
        const uint32_t h = n ? rng() % n : 0;
        std::stable_sort(a.begin(), a.begin() + h, byKey);
        std::stable_sort(a.begin() + h, a.end(), byKey);
        break;
    }
    case 3: std::reverse(a.begin(), a.end()); break;
    default: break;
    }
    for (uint32_t i = 0; i < n; ++i) a[i].v = i;
    return a;
}
}  // namespace

int main() {
    std::mt19937 rng(12345);
    long bad = 0, cases = 0;
    for (int it = 0; it < 6200; ++it) {
        const uint32_t n = rng() % (it < 4000 ? 200 : it < 6000 ? 4000 : 32768);
        const std::vector<E> a = randomCase(rng, n);
        bad += check(a, false);
        bad += check(a, true);
        bad += checkLevels(a, {0u, n});
        cases += 3;
    }
    // Several segments (worlds) at once, each sorted as its own std::sort.
    for (int it = 0; it < 400; ++it) {
        const uint32_t worlds = 1 + rng() % 40;
        std::vector<uint32_t> starts{0};
        std::vector<E> a;
        for (uint32_t w = 0; w < worlds; ++w) {
            std::vector<E> part = randomCase(rng, rng() % (it % 4 == 0 ? 3000 : 300));
            for (E& e : part) e.v += (uint32_t)a.size();
            a.insert(a.end(), part.begin(), part.end());
            starts.push_back((uint32_t)a.size());
        }
        bad += checkLevels(a, starts);
        cases += 1;
    }
    std::printf("cases %ld mismatches %ld\n", cases, bad);
    return bad != 0;
}
