// Order of the primary CSFs: the reference's selection sort (stda.f:651-668;
// stda-rw_dual.f:836-853) reproduced exactly, ties included.  At step i it
// takes the first position k >= i holding the minimum of key[i..n) (only a
// strictly smaller value replaces the candidate) and swaps positions i and k,
// which moves the element at i behind later equal keys: the result is not a
// stable sort when keys repeat.  A segment tree over (value, first index)
// finds k in O(log n), so the whole order costs O(n log n).
#pragma once

#include <cstddef>
#include <cstdint>
#include <utility>
#include <vector>

namespace stda {

inline void reference_selection_order(std::vector<double> key, std::vector<int32_t> &items) {
    const size_t n = key.size();
    if (n < 2) return;
    size_t sz = 1;
    while (sz < n) sz <<= 1;
    // node t holds the index of the first minimum of its range (n = empty)
    std::vector<size_t> tr(2 * sz, n);
    auto better = [&](size_t a, size_t b) {   // a left of b
        if (a == n) return b;
        if (b == n) return a;
        return key[b] < key[a] ? b : a;
    };
    for (size_t i = 0; i < n; ++i) tr[sz + i] = i;
    for (size_t t = sz - 1; t >= 1; --t) tr[t] = better(tr[2 * t], tr[2 * t + 1]);
    auto update = [&](size_t i) {
        for (size_t t = (sz + i) >> 1; t >= 1; t >>= 1) tr[t] = better(tr[2 * t], tr[2 * t + 1]);
    };
    auto query = [&](size_t lo) {   // first minimum of [lo, n)
        size_t l = lo + sz, r = n + sz, left = n, right = n;
        while (l < r) {
            if (l & 1) left = better(left, tr[l++]);
            if (r & 1) right = better(tr[--r], right);
            l >>= 1;
            r >>= 1;
        }
        return better(left, right);
    };
    for (size_t i = 0; i + 1 < n; ++i) {
        const size_t k = query(i);
        if (k == i || k == n || !(key[k] < key[i])) continue;
        std::swap(key[i], key[k]);
        std::swap(items[i], items[k]);
        update(i);
        update(k);
    }
}

}  // namespace stda
