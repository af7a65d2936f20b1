// Host check of renzoku_amd/csrc/par_sort.h: the block-parallel std::sort
// (run here lane by lane) against the one-lane restatement in std_order.h and,
// at the default depth limit, against std::sort of the toolchain's libstdc++
// itself: the same permutation, equal keys included.  Inputs: random keys with
// few to many distinct values, sorted / reversed / organ-pipe / constant runs,
// sizes around the 16-element threshold and up to 4096; and small forced depth
// limits, which send ranges to the heap-sort fallback at every level.  The host
// executor runs the lanes of each phase in a shuffled order, so a phase that
// depended on another lane's work in the same phase would show up here.
// Host code only: compiled by hipcc, runs without a GPU.
#include <algorithm>
#include <cstdio>
#include <cstring>
#include <numeric>
#include <random>
#include <utility>
#include <vector>

#include "par_sort.h"

using namespace rzk;

constexpr int NL = 256;
static int fails = 0;
static long checks = 0;

static bool pair_less(const std::pair<int, float>& a, const std::pair<int, float>& b) { return a.second < b.second; }

static void check(const std::vector<float>& keys, int depth0, const char* what) {
    const int n = (int)keys.size();
    std::vector<float> k1(keys), k2(keys);
    std::vector<int> v1(n), v2(n);
    std::iota(v1.begin(), v1.end(), 0);
    std::iota(v2.begin(), v2.end(), 0);
    stdord::sort_pairs(k1.data(), v1.data(), n, depth0);
    if (depth0 < 0) {
        std::vector<std::pair<int, float>> line(n);
        for (int i = 0; i < n; ++i) line[i] = {i, keys[i]};
        std::sort(line.begin(), line.end(), pair_less);
        for (int i = 0; i < n; ++i)
            if (line[i].first != v1[i]) {
                if (fails++ < 10) std::printf("one-lane sort differs from std::sort (%s, n=%d) at %d\n", what, n, i);
                return;
            }
    }
    std::vector<unsigned short> posL(n + 1), posR(n + 1), rid(n + 1);
    std::vector<unsigned char> fl(n + 1);
    std::vector<int> seg(18 * NL), la(NL), lb(NL), ws(2 * NL / 64 + 1);
    psort::Work<float, int> w{k2.data(), v2.data(), posL.data(), posR.data(), fl.data(), rid.data(), seg.data(),
                              la.data(), lb.data(), ws.data()};
    psort::sort_pairs<psort::HostLanes<NL>>(w, n, depth0);
    ++checks;
    for (int i = 0; i < n; ++i)
        if (v1[i] != v2[i] || std::memcmp(&k1[i], &k2[i], 4) != 0) {
            if (fails++ < 10)
                std::printf("parallel sort differs (%s, n=%d, depth0=%d) at %d: %d vs %d\n", what, n, depth0, i, v1[i],
                            v2[i]);
            return;
        }
}

int main() {
    std::mt19937 g(2013);
    const int sizes[] = {0, 1, 2, 3, 5, 15, 16, 17, 18, 19, 31, 32, 33, 34, 50, 63, 64, 65, 100, 255, 256, 257,
                         511, 1000, 1519, 2047, 2048, 3001, 4095, 4096};
    for (int n : sizes) {
        for (int levels : {1, 2, 3, 7, 50, 100000}) {
            for (int rep = 0; rep < 4; ++rep) {
                std::uniform_int_distribution<int> lv(0, levels - 1);
                std::vector<float> k(n);
                for (auto& x : k) x = (float)lv(g) * 0.25f;
                check(k, -1, "random");
                if (n > 16 && rep == 0)
                    for (int d0 : {0, 1, 2, 3, 5}) check(k, d0, "forced depth");
            }
        }
        std::vector<float> k(n);
        for (int i = 0; i < n; ++i) k[i] = (float)i;
        check(k, -1, "sorted");
        std::reverse(k.begin(), k.end());
        check(k, -1, "reversed");
        for (int i = 0; i < n; ++i) k[i] = (float)std::min(i, n - 1 - i);
        check(k, -1, "organ pipe");
        if (n > 16) check(k, 2, "organ pipe, forced depth");
        std::fill(k.begin(), k.end(), 1.5f);
        check(k, -1, "constant");
        for (int i = 0; i < n; ++i) k[i] = (i % 3 == 0) ? 0.0f : (float)(i % 17);
        check(k, -1, "zeros and a few values");
    }
    std::printf("%ld sorts, %d mismatches\n", checks, fails);
    return fails ? 1 : 0;
}
