// Host check of renzoku_amd/csrc/std_order.h against the standard library it
// restates (libstdc++ of this g++/clang toolchain): std::sort of (index, key)
// pairs by key and std::priority_queue of (index, cost) by cost, on inputs full
// of equal keys (where only the algorithms' own moves decide the order), sorted
// and reversed runs, and sizes around the 16-element insertion threshold; the
// heap-sort fallback is checked against std::partial_sort on the whole range.
// Host code only: compiled by hipcc, runs without a GPU.
#include <algorithm>
#include <cstdio>
#include <queue>
#include <random>
#include <utility>
#include <vector>

#include "std_order.h"

using namespace rzk;

static int fails = 0;
static long checks = 0;

static bool pair_less(const std::pair<int, float>& a, const std::pair<int, float>& b) { return a.second < b.second; }

static void check_sort(const std::vector<float>& keys, const char* what) {
    int n = (int)keys.size();
    std::vector<std::pair<int, float>> line(n);
    for (int i = 0; i < n; ++i) line[i] = {i, keys[i]};
    std::sort(line.begin(), line.end(), pair_less);
    std::vector<float> k(keys);
    std::vector<int> v(n);
    for (int i = 0; i < n; ++i) v[i] = i;
    stdord::sort_pairs(k.data(), v.data(), n);
    ++checks;
    for (int i = 0; i < n; ++i)
        if (line[i].first != v[i]) {
            if (fails++ < 10) std::printf("sort mismatch (%s, n=%d) at %d: %d vs %d\n", what, n, i, line[i].first, v[i]);
            return;
        }
}

static void check_heapsort(const std::vector<float>& keys) {
    int n = (int)keys.size();
    std::vector<std::pair<int, float>> line(n);
    for (int i = 0; i < n; ++i) line[i] = {i, keys[i]};
    std::partial_sort(line.begin(), line.end(), line.end(), pair_less);
    std::vector<float> k(keys);
    std::vector<int> v(n);
    for (int i = 0; i < n; ++i) v[i] = i;
    stdord::heap_sort_range(k.data(), v.data(), 0, n);
    ++checks;
    for (int i = 0; i < n; ++i)
        if (line[i].first != v[i]) {
            if (fails++ < 10) std::printf("heap sort mismatch (n=%d) at %d\n", n, i);
            return;
        }
}

struct CC { int index; float cost; };
struct CCLess { bool operator()(const CC& a, const CC& b) const { return a.cost < b.cost; } };

static void check_queue(std::mt19937& g, int levels, int ops) {
    std::priority_queue<CC, std::vector<CC>, CCLess> q;
    std::vector<float> key(ops + 1);
    std::vector<int> val(ops + 1);
    int n = 0, next = 0;
    std::uniform_int_distribution<int> lv(0, levels - 1), coin(0, 2);
    for (int t = 0; t < ops; ++t) {
        if (n > 0 && coin(g) == 0) {
            ++checks;
            if (q.top().index != val[0] || q.top().cost != key[0]) {
                if (fails++ < 10) std::printf("queue top mismatch at op %d\n", t);
                return;
            }
            q.pop();
            stdord::heap_pop(key.data(), val.data(), n);
        } else {
            float c = (float)lv(g);
            q.push({next, c});
            stdord::heap_push(key.data(), val.data(), n, c, next);
            ++next;
        }
    }
}

int main() {
    std::mt19937 g(2013);
    for (int n : {0, 1, 2, 3, 5, 15, 16, 17, 18, 31, 32, 33, 63, 64, 65, 100, 257, 1000, 4097}) {
        for (int levels : {1, 2, 3, 7, 50, 100000}) {
            for (int rep = 0; rep < 6; ++rep) {
                std::uniform_int_distribution<int> lv(0, levels - 1);
                std::vector<float> keys(n);
                for (auto& x : keys) x = (float)lv(g) * 0.25f - 3.0f;
                check_sort(keys, "random");
                check_heapsort(keys);
                std::vector<float> s(keys);
                std::sort(s.begin(), s.end());
                check_sort(s, "sorted");
                std::reverse(s.begin(), s.end());
                check_sort(s, "reversed");
            }
        }
    }
    // a pipe-organ shape and a sawtooth, which drive median-of-three pivots badly
    for (int n : {200, 2000, 20000}) {
        std::vector<float> a(n), b(n);
        for (int i = 0; i < n; ++i) { a[i] = (float)std::min(i, n - 1 - i); b[i] = (float)(i % 17); }
        check_sort(a, "organ");
        check_sort(b, "saw");
        check_heapsort(a);
    }
    for (int levels : {1, 2, 5, 1000})
        for (int rep = 0; rep < 20; ++rep) check_queue(g, levels, 3000);
    std::printf("std_order: %ld checks, %d mismatches\n", checks, fails);
    return fails == 0 ? 0 : 1;
}
