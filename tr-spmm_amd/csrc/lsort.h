// Emulation of GNU libstdc++'s std::sort (introsort + final insertion sort,
// as shipped in GCC 4.x through 14) for 32-bit elements under a caller-given
// strict weak order.  The reference router orders its 2:4 units with
// std::sort and a comparator that leaves ties unbroken
// (TR-source/Block/block_online.cpp:356-363), so the permutation it produces
// for equal keys is whatever libstdc++'s steps produce; reproducing the routed
// metadata bit-exactly means reproducing those steps:
//   introsort loop   while a range has > 16 elements: if the depth budget
//                    (2*floor(log2 n)) is spent, heap-sort the range;
//                    else move the median of (first+1, mid, last-1) to
//                    first, Hoare-partition [first+1, last) around it,
//                    recurse on the right part, loop on the left part;
//   final pass       guarded insertion sort of the first 16 elements and
//                    unguarded linear insertion of the rest.
// Everything is sequential and runs on one lane (device) or the host; the
// element array may live in LDS or global memory.
#pragma once

#include <cstdint>

#if defined(__HIPCC__)
#define LS_FN __host__ __device__ inline
#else
#define LS_FN inline
#endif

namespace trspmm {
namespace lsort {

constexpr int kThreshold = 16;

LS_FN int floor_log2(int64_t n) {
    int r = -1;
    while (n) {
        n >>= 1;
        ++r;
    }
    return r;
}

LS_FN void swap_at(uint32_t* a, int64_t i, int64_t j) {
    const uint32_t t = a[i];
    a[i] = a[j];
    a[j] = t;
}

// sift the hole at `hole` down to a leaf, then push `value` back up (heap of length len)
template <typename Cmp>
LS_FN void heap_adjust(uint32_t* a, int64_t hole, int64_t len, uint32_t value, Cmp less) {
    const int64_t top = hole;
    int64_t child = hole;
    while (child < (len - 1) / 2) {
        child = 2 * (child + 1);
        if (less(a[child], a[child - 1])) --child;
        a[hole] = a[child];
        hole = child;
    }
    if ((len & 1) == 0 && child == (len - 2) / 2) {
        child = 2 * (child + 1);
        a[hole] = a[child - 1];
        hole = child - 1;
    }
    int64_t parent = (hole - 1) / 2;
    while (hole > top && less(a[parent], value)) {
        a[hole] = a[parent];
        hole = parent;
        parent = (hole - 1) / 2;
    }
    a[hole] = value;
}

template <typename Cmp>
LS_FN void heap_sort(uint32_t* a, int64_t n, Cmp less) {
    if (n >= 2) {
        for (int64_t parent = (n - 2) / 2;; --parent) {
            heap_adjust(a, parent, n, a[parent], less);
            if (parent == 0) break;
        }
    }
    for (int64_t last = n; last > 1;) {
        --last;
        const uint32_t v = a[last];
        a[last] = a[0];
        heap_adjust(a, 0, last, v, less);
    }
}

template <typename Cmp>
LS_FN int64_t partition_pivot(uint32_t* a, int64_t first, int64_t last, Cmp less) {
    const int64_t mid = first + (last - first) / 2;
    const int64_t x = first + 1, y = mid, z = last - 1;
    int64_t m;
    if (less(a[x], a[y])) {
        if (less(a[y], a[z])) m = y;
        else if (less(a[x], a[z])) m = z;
        else m = x;
    } else if (less(a[x], a[z])) {
        m = x;
    } else if (less(a[y], a[z])) {
        m = z;
    } else {
        m = y;
    }
    swap_at(a, first, m);
    const uint32_t pivot = a[first];
    int64_t lo = first + 1, hi = last;
    while (true) {
        while (less(a[lo], pivot)) ++lo;
        --hi;
        while (less(pivot, a[hi])) --hi;
        if (!(lo < hi)) return lo;
        swap_at(a, lo, hi);
        ++lo;
    }
}

template <typename Cmp>
LS_FN void linear_insert(uint32_t* a, int64_t pos, Cmp less) {
    const uint32_t v = a[pos];
    int64_t next = pos - 1;
    while (less(v, a[next])) {
        a[pos] = a[next];
        pos = next;
        --next;
    }
    a[pos] = v;
}

template <typename Cmp>
LS_FN void insertion_sort(uint32_t* a, int64_t first, int64_t last, Cmp less) {
    if (first == last) return;
    for (int64_t i = first + 1; i != last; ++i) {
        if (less(a[i], a[first])) {
            const uint32_t v = a[i];
            for (int64_t k = i; k > first; --k) a[k] = a[k - 1];
            a[first] = v;
        } else {
            linear_insert(a, i, less);
        }
    }
}

// Sorts a[0..n) exactly as libstdc++'s std::sort(a, a + n, less) does.
// `stack` needs room for 2 * floor(log2 n) + 2 (first, last, depth) triples
// (pending ranges have strictly decreasing depth budgets, so at most
// depth + 1 are pending); n < 2^31;
// the recursion of the reference is unrolled onto it (ranges are disjoint, so
// the processing order does not change the result).
template <typename Cmp>
LS_FN void sort(uint32_t* a, int64_t n, Cmp less, int32_t* stack) {
    if (n <= 0) return;
    int sp = 0;
    stack[0] = 0;
    stack[1] = (int32_t)n;
    stack[2] = 2 * floor_log2(n);
    sp = 1;
    while (sp > 0) {
        --sp;
        int64_t first = stack[3 * sp], last = stack[3 * sp + 1];
        int64_t depth = stack[3 * sp + 2];
        while (last - first > kThreshold) {
            if (depth == 0) {
                heap_sort(a + first, last - first, less);
                break;
            }
            --depth;
            const int64_t cut = partition_pivot(a, first, last, less);
            stack[3 * sp] = (int32_t)cut;
            stack[3 * sp + 1] = (int32_t)last;
            stack[3 * sp + 2] = (int32_t)depth;
            ++sp;
            last = cut;
        }
    }
    if (n > kThreshold) {
        insertion_sort(a, 0, kThreshold, less);
        for (int64_t i = kThreshold; i < n; ++i) linear_insert(a, i, less);
    } else {
        insertion_sort(a, 0, n, less);
    }
}

// the stack depth sort() needs for n elements (in triples)
LS_FN int stack_triples(int64_t n) { return 2 * (n > 1 ? floor_log2(n) : 1) + 2; }

// comparator of the router: elements pack (key << 24) | index; order by key desc
struct KeyDesc {
    LS_FN bool operator()(uint32_t x, uint32_t y) const { return (x >> 24) > (y >> 24); }
};

}  // namespace lsort
}  // namespace trspmm
