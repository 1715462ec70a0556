// Device prefix sums and bounded searches used by the router and the plan
// builder.  Arrays up to kScanSingle entries are scanned by one workgroup
// (each thread walks its own contiguous segment: fine for short arrays, but
// uncoalesced, about 450 us for 2^18 entries); longer ones (per-row counts,
// per-nonzero flags) in three passes over chunks of kScanChunk entries: chunk
// sums, a one-workgroup scan of the sums, and each chunk's scan from its
// offset, with stream-ordered scratch for the chunk sums.
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

constexpr int64_t kScanSingle = 1 << 14;
constexpr int64_t kScanChunk = 1 << 14;

// sums[b] = sum of chunk b of `in` (null = zeros)
template <typename T>
__global__ __launch_bounds__(1024) void k_scan_chunk_sums(const T* __restrict__ in, int64_t n, T* __restrict__ sums) {
    __shared__ T part[1024];
    const int64_t b = (int64_t)blockIdx.x * kScanChunk, e = b + kScanChunk < n ? b + kScanChunk : n;
    T v = 0;
    if (in)
        for (int64_t i = b + threadIdx.x; i < e; i += 1024) v += in[i];
    part[threadIdx.x] = v;
    __syncthreads();
    for (int off = 512; off > 0; off >>= 1) {
        if ((int)threadIdx.x < off) part[threadIdx.x] += part[threadIdx.x + off];
        __syncthreads();
    }
    if (threadIdx.x == 0) sums[blockIdx.x] = part[0];
}

// out[b .. e) = exclusive scan of chunk blockIdx.x, starting from offs[blockIdx.x];
// out[n] = offs[gridDim.x], the total
template <typename T>
__global__ __launch_bounds__(1024) void k_scan_chunks(const T* __restrict__ in, T* __restrict__ out, int64_t n,
                                                      const T* __restrict__ offs) {
    constexpr int PER = (int)(kScanChunk / 1024);
    __shared__ T part[1024];
    const int tid = threadIdx.x;
    const int64_t b = (int64_t)blockIdx.x * kScanChunk + (int64_t)tid * PER;
    T v[PER];
    T s = 0;
#pragma unroll
    for (int k = 0; k < PER; ++k) {
        v[k] = (in && b + k < n) ? in[b + k] : T(0);
        s += v[k];
    }
    part[tid] = s;
    __syncthreads();
    for (int off = 1; off < 1024; off <<= 1) {
        const T x = tid >= off ? part[tid - off] : T(0);
        __syncthreads();
        part[tid] += x;
        __syncthreads();
    }
    T run = offs[blockIdx.x] + (tid ? part[tid - 1] : T(0));
#pragma unroll
    for (int k = 0; k < PER; ++k) {
        if (b + k < n) out[b + k] = run;
        run += v[k];
    }
    if (blockIdx.x == 0 && tid == 0) out[n] = offs[gridDim.x];     // offs has one entry per chunk + the total
}

template <typename T>
inline int exclusive_scan(const T* in, T* out, int64_t n, hipStream_t s, StreamArena* arena = nullptr) {
    if (n <= kScanSingle) {
        hipLaunchKernelGGL(k_exclusive_scan<T>, dim3(1), dim3(1024), 0, s, in, out, n);
        TRSPMM_HIP(hipGetLastError());
        return TRSPMM_OK;
    }
    const int64_t G = (n + kScanChunk - 1) / kScanChunk;
    T* chunk_sums = nullptr;
    if (arena) {
        DevBuf sums;          // back to the arena at once: later work on s reuses it in stream order
        if (int rc = sums.alloc(*arena, sizeof(T) * (size_t)(2 * G + 1))) return rc;
        T* cs = sums.as<T>();
        hipLaunchKernelGGL(k_scan_chunk_sums<T>, dim3((unsigned)G), dim3(1024), 0, s, in, n, cs);
        hipLaunchKernelGGL(k_exclusive_scan<T>, dim3(1), dim3(1024), 0, s, cs, cs + G, G);
        hipLaunchKernelGGL(k_scan_chunks<T>, dim3((unsigned)G), dim3(1024), 0, s, in, out, n, cs + G);
        TRSPMM_HIP(hipGetLastError());
        return TRSPMM_OK;
    }
    if (hipMallocAsync((void**)&chunk_sums, sizeof(T) * (size_t)(2 * G + 1), s) != hipSuccess) {
        (void)hipGetLastError();
        set_error("exclusive_scan: stream-ordered allocation of " + std::to_string(2 * G + 1) + " chunk sums failed");
        return TRSPMM_ENOMEM;
    }
    T* chunk_offs = chunk_sums + G;
    hipLaunchKernelGGL(k_scan_chunk_sums<T>, dim3((unsigned)G), dim3(1024), 0, s, in, n, chunk_sums);
    hipLaunchKernelGGL(k_exclusive_scan<T>, dim3(1), dim3(1024), 0, s, chunk_sums, chunk_offs, G);
    hipLaunchKernelGGL(k_scan_chunks<T>, dim3((unsigned)G), dim3(1024), 0, s, in, out, n, chunk_offs);
    const hipError_t launched = hipGetLastError();
    TRSPMM_HIP(hipFreeAsync(chunk_sums, s));          // released once the three kernels have run
    TRSPMM_HIP(launched);
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
