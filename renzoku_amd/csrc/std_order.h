// std_order.h — the element orders of two C++ standard-library algorithms the
// reference's LightSlice refinement depends on, restated for host and device:
//
//  * sort_pairs: std::sort (libstdc++, GCC 11: introsort with a median-of-three
//    pivot moved to the front, an unguarded Hoare partition, a depth limit of
//    2*floor(log2 n) after which the range is heap-sorted, and a final
//    insertion sort with a 16-element threshold) of (key, value) pairs by key,
//    used as std::sort(line.begin(), line.end(), pair_less) at
//    lightslice.cc:463.  The algorithm is not stable, so the order of equal
//    keys is whatever this sequence of moves leaves; following the same moves
//    gives the same order.
//  * heap_push / heap_pop: std::priority_queue<ClusterCost, vector,
//    ClusterCostLess> (lightslice.cc:379-383, :862-902): push_back + push_heap,
//    pop_heap + pop_back, a max-heap by cost whose top among equal costs is
//    again fixed by the sift moves.
//
// Keys are compared with operator< only, as the reference's comparators do.
#pragma once

#include "rzk_math.h"

namespace rzk {
namespace stdord {

// ---- heap primitives (the sift-down "hole" method), on parallel arrays -----
template <typename K, typename V>
RZK_HD void push_heap_hole(K* key, V* val, int hole, int top, K vk, V vv) {
    int parent = (hole - 1) / 2;
    while (hole > top && key[parent] < vk) {
        key[hole] = key[parent];
        val[hole] = val[parent];
        hole = parent;
        parent = (hole - 1) / 2;
    }
    key[hole] = vk;
    val[hole] = vv;
}

template <typename K, typename V>
RZK_HD void adjust_heap(K* key, V* val, int hole, int len, K vk, V vv) {
    const int top = hole;
    int child = hole;
    while (child < (len - 1) / 2) {
        child = 2 * (child + 1);
        if (key[child] < key[child - 1]) child--;
        key[hole] = key[child];
        val[hole] = val[child];
        hole = child;
    }
    if ((len & 1) == 0 && child == (len - 2) / 2) {
        child = 2 * (child + 1);
        key[hole] = key[child - 1];
        val[hole] = val[child - 1];
        hole = child - 1;
    }
    push_heap_hole(key, val, hole, top, vk, vv);
}

// priority_queue::push: the new element goes to position n, then sifts up
template <typename K, typename V>
RZK_HD void heap_push(K* key, V* val, int& n, K k, V v) {
    key[n] = k;
    val[n] = v;
    ++n;
    push_heap_hole(key, val, n - 1, 0, k, v);
}

// priority_queue::pop: the last element is sifted down from the top
template <typename K, typename V>
RZK_HD void heap_pop(K* key, V* val, int& n) {
    if (n > 1) {
        K lk = key[n - 1];
        V lv = val[n - 1];
        key[n - 1] = key[0];
        val[n - 1] = val[0];
        adjust_heap(key, val, 0, n - 1, lk, lv);
    }
    --n;
}

// ---- std::sort ---------------------------------------------------------------
template <typename K, typename V>
RZK_HD void swap_at(K* key, V* val, int a, int b) {
    K tk = key[a];
    key[a] = key[b];
    key[b] = tk;
    V tv = val[a];
    val[a] = val[b];
    val[b] = tv;
}

// heap sort of [first, last) (std::__partial_sort with middle == last)
template <typename K, typename V>
RZK_HD void heap_sort_range(K* key, V* val, int first, int last) {
    K* k = key + first;
    V* v = val + first;
    int len = last - first;
    if (len >= 2) {                              // make_heap
        for (int parent = (len - 2) / 2;; --parent) {
            adjust_heap(k, v, parent, len, k[parent], v[parent]);
            if (parent == 0) break;
        }
    }
    while (len > 1) {                            // sort_heap
        --len;
        K lk = k[len];
        V lv = v[len];
        k[len] = k[0];
        v[len] = v[0];
        adjust_heap(k, v, 0, len, lk, lv);
    }
}

// median of (a, b, c) moved to `result`
template <typename K, typename V>
RZK_HD void median_to_first(K* key, V* val, int result, int a, int b, int c) {
    if (key[a] < key[b]) {
        if (key[b] < key[c]) swap_at(key, val, result, b);
        else if (key[a] < key[c]) swap_at(key, val, result, c);
        else swap_at(key, val, result, a);
    } else if (key[a] < key[c]) {
        swap_at(key, val, result, a);
    } else if (key[b] < key[c]) {
        swap_at(key, val, result, c);
    } else {
        swap_at(key, val, result, b);
    }
}

template <typename K, typename V>
RZK_HD int unguarded_partition(K* key, V* val, int first, int last, int pivot) {
    for (;;) {
        while (key[first] < key[pivot]) ++first;
        --last;
        while (key[pivot] < key[last]) --last;
        if (!(first < last)) return first;
        swap_at(key, val, first, last);
        ++first;
    }
}

template <typename K, typename V>
RZK_HD void unguarded_linear_insert(K* key, V* val, int last) {
    K vk = key[last];
    V vv = val[last];
    int next = last - 1;
    while (vk < key[next]) {
        key[last] = key[next];
        val[last] = val[next];
        last = next;
        --next;
    }
    key[last] = vk;
    val[last] = vv;
}

template <typename K, typename V>
RZK_HD void insertion_sort(K* key, V* val, int first, int last) {
    if (first == last) return;
    for (int i = first + 1; i != last; ++i) {
        if (key[i] < key[first]) {
            K vk = key[i];
            V vv = val[i];
            for (int j = i; j > first; --j) {
                key[j] = key[j - 1];
                val[j] = val[j - 1];
            }
            key[first] = vk;
            val[first] = vv;
        } else {
            unguarded_linear_insert(key, val, i);
        }
    }
}

constexpr int kSortThreshold = 16;
constexpr int kSortStack = 64;     // pending ranges: at most one per depth level (depth <= 2 * 31)

RZK_HD int floor_log2(int n) {
    int r = 0;
    while (n > 1) { n >>= 1; ++r; }
    return r;
}

// std::sort(key/val[0, n)) by key.  The reference's recursion on the right
// part is replaced by a stack of pending ranges: the parts are disjoint, so
// the order they are finished in does not change any element's final place.
// depth0 < 0: the introsort depth limit 2*floor(log2 n) (tests pass a smaller
// one to reach the heap-sort fallback).
template <typename K, typename V>
RZK_HD void sort_pairs(K* key, V* val, int n, int depth0 = -1) {
    if (n <= 1) return;
    int st_first[kSortStack], st_last[kSortStack], st_depth[kSortStack];
    int sp = 0;
    st_first[sp] = 0;
    st_last[sp] = n;
    st_depth[sp] = depth0 >= 0 ? depth0 : 2 * floor_log2(n);
    ++sp;
    while (sp > 0) {
        --sp;
        int first = st_first[sp], last = st_last[sp], depth = st_depth[sp];
        while (last - first > kSortThreshold) {
            if (depth == 0) {
                heap_sort_range(key, val, first, last);
                break;
            }
            --depth;
            int mid = first + (last - first) / 2;
            median_to_first(key, val, first, first + 1, mid, last - 1);
            int cut = unguarded_partition(key, val, first + 1, last, first);
            st_first[sp] = cut;
            st_last[sp] = last;
            st_depth[sp] = depth;
            ++sp;
            last = cut;
        }
    }
    if (n > kSortThreshold) {
        insertion_sort(key, val, 0, kSortThreshold);
        for (int i = kSortThreshold; i < n; ++i) unguarded_linear_insert(key, val, i);
    } else {
        insertion_sort(key, val, 0, n);
    }
}

}  // namespace stdord
}  // namespace rzk
