// Device prefix sums and bounded searches used by the router and the plan
// builder.  One workgroup scans the whole array: every caller scans per-window
// or per-part counts (at most a few million entries), where a single pass of
// 1024 threads is a few tens of microseconds and needs no scratch memory.
#pragma once

#include "common.h"

namespace trspmm {

// out[i] = sum(in[0..i)), i = 0..n; out has n + 1 entries (out[n] = total).
// `in` may be null (treated as zeros) — used to zero-fill offsets.
template <typename T>
__global__ __launch_bounds__(1024) void k_exclusive_scan(const T* __restrict__ in, T* __restrict__ out,
                                                         int64_t n) {
    __shared__ T part[1024];
    const int tid = threadIdx.x;
    const int64_t per = (n + 1023) / 1024;
    const int64_t b = (int64_t)tid * per, e = b + per < n ? b + per : n;
    T s = 0;
    for (int64_t i = b; i < e; ++i) s += in ? in[i] : T(0);
    part[tid] = s;
    __syncthreads();
    // Hillis-Steele inclusive scan of the 1024 partial sums
    for (int off = 1; off < 1024; off <<= 1) {
        T v = tid >= off ? part[tid - off] : T(0);
        __syncthreads();
        part[tid] += v;
        __syncthreads();
    }
    T run = tid ? part[tid - 1] : T(0);
    for (int64_t i = b; i < e; ++i) {
        out[i] = run;
        run += in ? in[i] : T(0);
    }
    if (tid == 1023) out[n] = part[1023];
}

template <typename T>
inline int exclusive_scan(const T* in, T* out, int64_t n, hipStream_t s) {
    hipLaunchKernelGGL(k_exclusive_scan<T>, dim3(1), dim3(1024), 0, s, in, out, n);
    TRSPMM_HIP(hipGetLastError());
    return TRSPMM_OK;
}

// first index i in [0, n) with a[i] >= key (n if none); `a` non-decreasing
template <typename T>
__device__ inline int64_t lower_bound(const T* a, int64_t n, T key) {
    int64_t lo = 0, hi = n;
    while (lo < hi) {
        const int64_t mid = (lo + hi) >> 1;
        if (a[mid] < key) lo = mid + 1;
        else hi = mid;
    }
    return lo;
}

}  // namespace trspmm
