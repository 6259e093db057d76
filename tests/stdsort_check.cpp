// Checks the libstdc++ std::sort restatement (velox_amd/csrc/vx_stdsort.hpp)
// against std::sort and std::partial_sort of the host's libstdc++ on seeded
// random (key, index) sequences with many equal keys: the index order left
// among equal keys must be identical.  Prints the mismatch count.
#include "../velox_amd/csrc/vx_stdsort.hpp"
#include <algorithm>
#include <cstdio>
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
}  // namespace

int main() {
    std::mt19937 rng(12345);
    long bad = 0, cases = 0;
    for (int it = 0; it < 6000; ++it) {
        const uint32_t n = rng() % (it < 4000 ? 200 : 4000);
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
        bad += check(a, false);
        bad += check(a, true);
        cases += 2;
    }
    std::printf("cases %ld mismatches %ld\n", cases, bad);
    return bad != 0;
}
