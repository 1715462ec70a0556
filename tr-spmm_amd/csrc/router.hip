// Device router: canonical CSR -> the reference's 29 routed-metadata arrays,
// bit-exact with Libra5BlockOnline.block_sptc_2to4_online
// (TR-source/Block/block_online.cpp:545-871) for window <= 16.
//
// For window <= 16 every per-row field of a column's 2-bit row vector is 0
// or 1 (block_online.cpp:592-600), so the 2:4 compatibility test
// (block_online.cpp:85-88) always passes and the reference's unit building
// and max-degree matching reduce to (per window):
//   units  = consecutive pairs of the non-TC columns ordered by
//            (count desc, column asc)                  (:267-328)
//   order  = libstdc++ std::sort of iota(units) by unit nnz desc (:356-363)
//   pairs  = each unmatched unit in that order takes the lowest-index
//            unmatched unit (:365-386: all degrees equal and unit nnz is
//            non-increasing in the index)
// The oracle checks this closed form against the step-by-step restatement
// and against the reference router itself.
//
// Pipeline (all on the caller's stream; one host sync to size the outputs):
//   1. validate the CSR; build keys (window << colbits | column) per edge
//   2. stable radix sort by key (rocPRIM): each (window, column) becomes a run
//      of its edges in row order; run-length encode
//   3. one wave per window: stable counting sort of its runs by count desc,
//      TC chunk selection, unit list, libstdc++ order emulation (one lane for
//      small windows, the whole wave beyond 256 units; in LDS when it fits),
//      pairing (the wave: a unit is taken iff its index is at most the latest
//      partner, and a partner search is one ballot)
//   4. prefix sums over windows -> output sizes; allocate outputs
//   5. emission kernels: TC sub-blocks/parts, 2:4 groups and per-row slots,
//      residual flags -> per-row residual parts, mma.sp tile packing.
#include <chrono>
#include <cstring>   // rocPRIM's texture iterator uses memset without including it
#include <algorithm>
#include <array>
#include <map>
#include <exception>
#include <mutex>
#include <string>
#include <thread>
#include <vector>

#include <rocprim/rocprim.hpp>

#include "lsort.h"
#include "router.h"
#include "scan.h"

namespace trspmm {

namespace {

constexpr int kWave = 64;
constexpr int kWaveSortMin = 256;   // windows with more 2:4 units sort them with the whole wave

__device__ inline uint64_t lanemask_lt() {
    const int l = threadIdx.x & 63;
    return l ? (~0ull >> (64 - l)) : 0ull;
}

__device__ inline uint64_t ballot(bool p) { return __ballot(p); }

// window w with off[w] <= x < off[w + 1] (off non-decreasing, off[0] = 0)
__device__ inline int64_t upper_bound_window(const int32_t* off, int64_t W, int32_t x) {
    int64_t lo = 0, hi = W + 1;
    while (lo < hi) {
        const int64_t mid = (lo + hi) >> 1;
        if (off[mid] <= x) lo = mid + 1;
        else hi = mid;
    }
    return lo - 1;
}

// ---------------------------------------------------------------- 1. validation + keys
// One wave per row (four rows per 256-thread block): hub rows of tens of
// thousands of nonzeros would otherwise serialise on one thread.
__device__ inline int64_t wave_row() { return (int64_t)blockIdx.x * 4 + (threadIdx.x >> 6); }

__device__ inline int64_t wave_min_i64(int64_t v) {
    for (int off = 32; off > 0; off >>= 1) {
        const int64_t o = __shfl_xor(v, off);
        v = o < v ? o : v;
    }
    return v;
}

// err bit 1: bad row pointers; 2: a column outside [0, ncols); 4: columns not
// strictly ascending.  Within a row the first offending entry decides between
// 2 and 4 (range checked first), as a sequential scan would.
__global__ __launch_bounds__(256) void k_validate(const int32_t* __restrict__ rowptr, const int32_t* __restrict__ col,
                                                  int64_t rows, int64_t nnz, int64_t ncols, int* __restrict__ err) {
    const int64_t r = wave_row();
    const int lane = threadIdx.x & 63;
    if (r == 0 && lane == 0 && (rowptr[0] != 0 || (int64_t)rowptr[rows] != nnz)) atomicOr(err, 1);
    if (r >= rows) return;
    const int32_t b = rowptr[r], e = rowptr[r + 1];
    if (b > e || b < 0 || (int64_t)e > nnz) {
        if (lane == 0) atomicOr(err, 1);
        return;
    }
    int64_t first = INT64_MAX;                     // 4 * position + (2: range, 3: order), lowest position first
    for (int32_t i = b + lane; i < e && first == INT64_MAX; i += kWave) {
        const int32_t c = col[i];
        if (c < 0 || (int64_t)c >= ncols) first = 4 * (int64_t)i + 2;
        else if (i > b && c <= col[i - 1]) first = 4 * (int64_t)i + 3;
    }
    first = wave_min_i64(first);
    if (lane == 0 && first != INT64_MAX) atomicOr(err, (first & 3) == 2 ? 2 : 4);
}

__global__ void k_keys(const int32_t* __restrict__ rowptr, const int32_t* __restrict__ col, int64_t rows, int64_t nnz,
                       int window, int colbits, uint64_t* __restrict__ keys, int32_t* __restrict__ vals,
                       int32_t* __restrict__ erow) {
    for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < nnz; e += (int64_t)gridDim.x * blockDim.x) {
        // row = last r with rowptr[r] <= e
        int64_t lo = 0, hi = rows;
        while (lo < hi) {
            const int64_t mid = (lo + hi + 1) >> 1;
            if (rowptr[mid] <= e) lo = mid;
            else hi = mid - 1;
        }
        const int64_t w = lo / window;
        keys[e] = ((uint64_t)w << colbits) | (uint64_t)col[e];
        vals[e] = (int32_t)e;
        erow[e] = (int32_t)lo;
    }
}

// ---------------------------------------------------------------- 2. runs
__global__ void k_run_flags(const uint64_t* __restrict__ keys, int64_t n, int32_t* __restrict__ flag) {
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
        flag[i] = (i == 0 || keys[i] != keys[i - 1]) ? 1 : 0;
}

__global__ void k_runs(const uint64_t* __restrict__ keys, const int32_t* __restrict__ flag,
                       const int32_t* __restrict__ run_of, int64_t n, int colbits, int32_t* __restrict__ run_start,
                       int32_t* __restrict__ run_col, int32_t* __restrict__ run_win) {
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
        if (!flag[i]) continue;
        const int32_t r = run_of[i];
        run_start[r] = (int32_t)i;
        run_col[r] = (int32_t)(keys[i] & ((1ull << colbits) - 1ull));
        run_win[r] = (int32_t)(keys[i] >> colbits);
    }
}

__global__ void k_win_runs(const int32_t* __restrict__ run_win, int64_t nruns, int64_t W, int32_t* __restrict__ win_rb) {
    const int64_t w = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (w <= W) win_rb[w] = (int32_t)lower_bound(run_win, nruns, (int32_t)w);
}

// ---------------------------------------------------------------- 3. per-window decisions
struct WinOut {
    int32_t* n_tcsub;     // TC sub-blocks (kept chunks)
    int32_t* n_tcval;     // TC nonzeros
    int32_t* n_pairs;     // matched unit pairs
    int32_t* n_groups;    // 2:4 groups (pairs + kept leftover)
    int32_t* leftover;    // unmatched unit index or -1
    int32_t* fallback;    // some column is in neither TC nor 2:4 (block_online.cpp:706-713)
    int32_t* n_rest;      // non-TC columns
};

struct RunArrays {
    const int32_t* start;   // [nruns + 1] first sorted-edge position
    const int32_t* col;
    const int32_t* win_rb;  // [W + 1]
    int32_t* order;         // [nruns] runs of a window by (count desc, col asc)
    int32_t* rest;          // [nruns] non-TC runs in that order
    int32_t* tc_chunk;      // [nruns] kept TC chunk indices
    uint8_t* in_tc;         // [nruns]
    uint32_t* units_g;      // [nruns] unit keys for windows whose units do not fit in LDS
    uint8_t* used_g;        // [nruns]
    int32_t* pair_a;        // [nruns]
    int32_t* pair_b;        // [nruns]
};

__device__ inline int run_count(const RunArrays& R, int32_t run) { return R.start[run + 1] - R.start[run]; }

// ---------------------------------------------------------------- wave-parallel std::sort emulation
//
// The same permutation as lsort::sort(a, n, lsort::KeyDesc{}) -- libstdc++'s
// introsort under a comparator that leaves ties unbroken -- computed by one
// wave.  The introsort loop visits the same ranges with the same pivots; each
// Hoare partition is computed in closed form.  With kp the pivot's key, the
// left scan stops at the positions of [first+1, last) whose key is <= kp
// (l_1 < l_2 < ...) and the right scan at those of [first, last) whose key is
// >= kp (r_1 > r_2 > ...; `first` holds the pivot).  Between two stops every
// position still holds its original value, so the partition swaps l_k with
// r_k for k = 1..S, S = max over x of min(#l <= x, #r > x), and returns
// min(l_{S+1}, r_S).  The final insertion pass moves an element only past
// strictly larger keys, so it is a stable sort by key: a counting sort
// through `tmp`.  A range whose depth budget runs out is heap-sorted by lane
// 0, as lsort::sort does.  `tmp` needs n entries of global memory.
struct PartLds {
    int32_t kp, cut_l;
};
struct WaveSortLds {
    int32_t base[256];     // per-key output offsets of the final pass
    PartLds part;
};

// Barriers of the sort: a one-wave workgroup uses the workgroup barrier; a
// wave of a larger workgroup working on a range of its own syncs only itself
// (its memory operations complete and become visible to its own lanes).
struct BlockSync {
    __device__ void operator()() const { __syncthreads(); }
};
struct WaveSync {
    __device__ void operator()() const {
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup");
    }
};

__device__ inline int32_t wave_max_i32(int32_t v) {
    for (int off = 32; off > 0; off >>= 1) v = max(v, __shfl_xor(v, off));
    return v;
}

template <typename Sync = BlockSync>
__device__ int32_t wave_partition(uint32_t* a, int32_t first, int32_t last, int32_t* __restrict__ tmp,
                                  PartLds& sl) {
    const Sync sync{};
    const int lane = threadIdx.x & (kWave - 1);
    if (lane == 0) {
        // median of (first + 1, mid, last - 1) moved to first (lsort::partition_pivot)
        const lsort::KeyDesc less{};
        const int32_t mid = first + (last - first) / 2, x = first + 1, y = mid, z = last - 1;
        int32_t m;
        if (less(a[x], a[y])) m = less(a[y], a[z]) ? y : (less(a[x], a[z]) ? z : x);
        else if (less(a[x], a[z])) m = x;
        else if (less(a[y], a[z])) m = z;
        else m = y;
        lsort::swap_at(a, first, m);
        sl.kp = (int32_t)(a[first] >> 24);
        sl.cut_l = INT32_MAX;
    }
    sync();
    const uint32_t kp = (uint32_t)sl.kp;
    const uint64_t le = lanemask_lt() | (1ull << lane);
    // #r over the range
    int32_t TR = 0;
    for (int32_t b = first; b < last; b += kWave) {
        const int32_t p = b + lane;
        TR += __popcll(ballot(p < last && (a[p] >> 24) >= kp));
    }
    // S = max_x min(#l <= x, #r > x): a running maximum per lane, one wave
    // reduction at the end
    int32_t PL = 0, PR = 0, S = 0;
    for (int32_t b = first; b < last; b += kWave) {
        const int32_t p = b + lane;
        const bool on = p < last;
        const uint32_t k = on ? a[p] >> 24 : 0u;
        const uint64_t mL = ballot(on && p > first && k <= kp), mR = ballot(on && k >= kp);
        const int32_t pl = PL + __popcll(mL & le), pr = PR + __popcll(mR & le);
        S = max(S, on ? min(pl, TR - pr) : 0);
        PL += __popcll(mL);
        PR += __popcll(mR);
    }
    S = wave_max_i32(S);
    // r_k (k <= S) -> tmp[k - 1]; l_{S+1} -> cut_l (nothing is swapped yet)
    PL = 0;
    PR = 0;
    for (int32_t b = first; b < last; b += kWave) {
        const int32_t p = b + lane;
        const bool on = p < last;
        const uint32_t k = on ? a[p] >> 24 : 0u;
        const bool isL = on && p > first && k <= kp, isR = on && k >= kp;
        const uint64_t mL = ballot(isL), mR = ballot(isR);
        const int32_t pl = PL + __popcll(mL & le), rank = TR - (PR + __popcll(mR & le)) + 1;
        if (isR && rank <= S) tmp[rank - 1] = p;
        if (isL && pl == S + 1) sl.cut_l = p;
        PL += __popcll(mL);
        PR += __popcll(mR);
    }
    sync();
    // swap l_k with r_k, k <= S.  Every swapped r lies above l_S, so the scan
    // ends after the chunk holding l_S, before reaching a swapped position
    // from an earlier chunk; within a chunk all reads precede the writes.
    PL = 0;
    for (int32_t b = first; b < last && PL < S; b += kWave) {
        const int32_t p = b + lane;
        const bool on = p < last;
        const uint32_t v = on ? a[p] : 0u;
        const bool isL = on && p > first && (v >> 24) <= kp;
        const uint64_t mL = ballot(isL);
        const int32_t pl = PL + __popcll(mL & le);
        if (isL && pl <= S) {
            const int32_t j = tmp[pl - 1];
            a[p] = a[j];
            a[j] = v;
        }
        PL += __popcll(mL);
    }
    const int32_t r_s = S > 0 ? tmp[S - 1] : INT32_MAX;
    int32_t cut = min(sl.cut_l, r_s);
    sync();
    return cut > last ? last : cut;     // a left stop always exists (libstdc++'s median-of-3 invariant)
}

// stable sort of a[0, n) by key descending, through tmp
__device__ void wave_stable_by_key(uint32_t* a, int32_t n, uint32_t* __restrict__ tmp, WaveSortLds& sl) {
    const int lane = threadIdx.x;
    const uint64_t lt = lanemask_lt();
    for (int t = lane; t < 256; t += kWave) sl.base[t] = 0;
    __syncthreads();
    for (int32_t p = lane; p < n; p += kWave) atomicAdd(&sl.base[a[p] >> 24], 1);
    __syncthreads();
    if (lane == 0) {
        int32_t acc = 0;
        for (int k = 255; k >= 0; --k) {
            const int32_t c = sl.base[k];
            sl.base[k] = acc;
            acc += c;
        }
    }
    __syncthreads();
    for (int32_t b = 0; b < n; b += kWave) {
        const int32_t p = b + lane;
        const bool on = p < n;
        const uint32_t v = on ? a[p] : 0u;
        const uint32_t k = v >> 24;
        uint64_t rem = ballot(on);
        while (rem) {
            const int leader = __ffsll((long long)rem) - 1;
            const uint32_t kk = (uint32_t)__shfl((int)k, leader);
            const uint64_t grp = ballot(on && k == kk);
            const int32_t at = sl.base[kk];          // read by every lane before lane 0 writes it
            if (on && k == kk) tmp[at + __popcll(grp & lt)] = v;
            if (lane == 0) sl.base[kk] = at + __popcll(grp);
            rem &= ~grp;
        }
    }
    __syncthreads();
    for (int32_t p = lane; p < n; p += kWave) a[p] = tmp[p];
    __syncthreads();
}

__device__ void wave_sort_desc(uint32_t* a, int32_t n, int32_t* __restrict__ tmp, int32_t* stack, WaveSortLds& sl) {
    const int lane = threadIdx.x;
    if (lane == 0) {
        stack[0] = 0;
        stack[1] = n;
        stack[2] = 2 * lsort::floor_log2(n);
    }
    __syncthreads();
    int sp = 1;
    while (sp > 0) {
        --sp;
        int32_t first = stack[3 * sp], last = stack[3 * sp + 1], depth = stack[3 * sp + 2];
        __syncthreads();                          // every lane has the entry before it is reused
        while (last - first > lsort::kThreshold) {
            if (depth == 0) {
                if (lane == 0) lsort::heap_sort(a + first, last - first, lsort::KeyDesc{});
                __syncthreads();
                break;
            }
            --depth;
            const int32_t cut = wave_partition(a, first, last, tmp, sl.part);
            if (lane == 0) {
                stack[3 * sp] = cut;
                stack[3 * sp + 1] = last;
                stack[3 * sp + 2] = depth;
            }
            __syncthreads();
            ++sp;
            last = cut;
        }
    }
    wave_stable_by_key(a, n, (uint32_t*)tmp, sl);
}

// One wave per window; the launch serves the windows of one size class,
// listed in `wlist` (largest first).  `lds_cap` = unit capacity of the
// dynamic LDS (0: units live in global memory).
__global__ __launch_bounds__(64) void k_window_decide(RunArrays R, WinOut O, int64_t W, int window, int density,
                                                      float tc_min_util, int skip_tc, int unit_min_nnz,
                                                      const int32_t* __restrict__ wlist, int lds_cap) {
    extern __shared__ uint32_t dyn[];
    __shared__ int32_t hist[17], base[17], fill[17];
    __shared__ int32_t s_ncand;
    __shared__ int32_t sort_stack[3 * 66];
    __shared__ WaveSortLds wsl;
    const int64_t w = wlist[blockIdx.x];
    const int lane = threadIdx.x;
    const int32_t rb = R.win_rb[w], re = R.win_rb[w + 1];
    const int32_t nc = re - rb;

    // a. histogram of run counts (1..16)
    if (lane < 17) {
        hist[lane] = 0;
        fill[lane] = 0;
    }
    __syncthreads();
    for (int32_t i = rb + lane; i < re; i += kWave) atomicAdd(&hist[run_count(R, i)], 1);
    __syncthreads();
    if (lane == 0) {
        int32_t acc = 0;
        for (int c = 16; c >= 1; --c) {
            base[c] = acc;
            acc += hist[c];
        }
        int32_t ncand = 0;
        if (!skip_tc)
            for (int c = 16; c >= 1 && c >= density; --c) ncand += hist[c];
        s_ncand = ncand;
    }
    __syncthreads();

    // b. stable counting sort by count desc (runs arrive in column order)
    for (int32_t i0 = rb; i0 < re; i0 += kWave) {
        const int32_t i = i0 + lane;
        const bool on = i < re;
        const int c = on ? run_count(R, i) : 0;
        const uint64_t lt = lanemask_lt();
        int32_t pos = 0;
        for (int v = 1; v <= 16; ++v) {
            const uint64_t m = ballot(on && c == v);
            if (on && c == v) pos = base[v] + fill[v] + __popcll(m & lt);
            __syncthreads();
            if (lane == 0) fill[v] += __popcll(m);
            __syncthreads();
        }
        if (on) R.order[rb + pos] = i;
    }
    __syncthreads();

    // c. TC chunks of 8 over the candidates (block_online.cpp:215-253)
    const int32_t ncand = s_ncand;
    const int32_t nchunk = (ncand + 7) / 8;
    int32_t kept_total = 0, tcval = 0;
    for (int32_t j0 = 0; j0 < nchunk; j0 += kWave) {
        const int32_t j = j0 + lane;
        bool keep = false;
        int32_t nnz = 0;
        if (j < nchunk) {
            const int32_t e = (8 * j + 8 < ncand) ? 8 * j + 8 : ncand;
            for (int32_t k = 8 * j; k < e; ++k) nnz += run_count(R, R.order[rb + k]);
            const float util = (float)nnz / (float)(window * 8);
            keep = !(util < tc_min_util);
            for (int32_t k = 8 * j; k < e; ++k) R.in_tc[R.order[rb + k]] = keep ? 1 : 0;
        }
        const uint64_t m = ballot(keep);
        if (keep) R.tc_chunk[rb + kept_total + __popcll(m & lanemask_lt())] = j;
        kept_total += __popcll(m);
        int32_t v = keep ? nnz : 0;
        for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off);
        tcval += v;
    }
    // candidates never reach the TC set when skipped; everything past ncand is not a candidate
    for (int32_t k = ncand + lane; k < nc; k += kWave) R.in_tc[R.order[rb + k]] = 0;
    __syncthreads();

    // d. non-TC runs in (count desc, col asc) order
    int32_t nrest = 0;
    for (int32_t k0 = 0; k0 < nc; k0 += kWave) {
        const int32_t k = k0 + lane;
        const int32_t run = k < nc ? R.order[rb + k] : 0;
        const bool on = k < nc && !R.in_tc[run];
        const uint64_t m = ballot(on);
        if (on) R.rest[rb + nrest + __popcll(m & lanemask_lt())] = run;
        nrest += __popcll(m);
    }
    __syncthreads();

    // e. units and their order key
    const int32_t nu = (nrest + 1) / 2;
    const bool in_lds = lds_cap > 0 && nu <= lds_cap;
    uint32_t* units = in_lds ? dyn : R.units_g + rb;
    for (int32_t k = lane; k < nu; k += kWave) {
        int32_t nnz = run_count(R, R.rest[rb + 2 * k]);
        if (2 * k + 1 < nrest) nnz += run_count(R, R.rest[rb + 2 * k + 1]);
        units[k] = ((uint32_t)nnz << 24) | (uint32_t)k;
    }
    uint8_t* used = in_lds ? (uint8_t*)(dyn + lds_cap) : R.used_g + rb;
    for (int32_t k = lane; k < nu; k += kWave) used[k] = 0;
    __syncthreads();

    // f. libstdc++ order (one wave for large windows, else one lane), then the
    // pairing of match_sptc_units_max_degree (block_online.cpp:365-391): in that
    // order, each unit not yet used takes the lowest-index unused unit as its
    // partner.  The partners are therefore taken in increasing index order, and
    // every index up to the latest partner `hw` is used (a partner, or a unit
    // that took one).  So a unit is used at its turn exactly when its index is
    // <= hw, and an active unit v's partner is the first index above hw, other
    // than v, that is not an earlier active unit.  used[] marks the active
    // units; a 64-lane ballot over used[hw + 1 ..] finds the partner, and the
    // order is read 64 units at a time into registers.
    const bool wave_sorted = nu > kWaveSortMin;
    if (wave_sorted) wave_sort_desc(units, nu, R.pair_a + rb, sort_stack, wsl);
    else if (lane == 0) lsort::sort(units, nu, lsort::KeyDesc{}, sort_stack);
    __syncthreads();
    // first index > from, other than `except`, not marked in used[] (nu if none)
    auto next_free = [&](int32_t from, int32_t except) -> int32_t {
        for (int32_t b = from + 1; b < nu; b += kWave) {
            const int32_t k = b + lane;
            const bool ok = (k < nu) & (k != except) && !used[k < nu ? k : 0];
            const uint64_t m = ballot(ok);
            if (m) return b + (int32_t)__builtin_ctzll(m);
        }
        return nu;
    };
    // pair np is held by lane np % 64 and stored with its 63 neighbours
    int32_t hw = -1, np = 0, my_a = 0, my_b = 0;
    for (int32_t i0 = 0; i0 < nu; i0 += kWave) {
        const int32_t mine = i0 + lane < nu ? (int32_t)(units[i0 + lane] & 0xffffffu) : 0;
        const int32_t cnt = nu - i0 < kWave ? nu - i0 : kWave;
        for (int32_t j = 0; j < cnt; ++j) {
            const int32_t v = __builtin_amdgcn_readlane(mine, j);
            if (v <= hw) continue;                       // already a partner
            const int32_t p = next_free(hw, v);
            if (p >= nu) continue;                       // no partner: v stays unused
            if (lane == 0) used[v] = 1;
            if (lane == (np & (kWave - 1))) {
                my_a = v;
                my_b = p;
            }
            hw = p;
            if ((++np & (kWave - 1)) == 0) {
                R.pair_a[rb + np - kWave + lane] = my_a;
                R.pair_b[rb + np - kWave + lane] = my_b;
            }
            // lane 0's mark is seen by the next search: shared memory serves one
            // wave in order; the global-memory flags of windows beyond the LDS
            // capacity need the stores done and the vector L1 refreshed
            if (in_lds) {
                __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
                __builtin_amdgcn_wave_barrier();
                __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
            } else {
                __builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent");
                __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
            }
        }
    }
    if (lane < (np & (kWave - 1))) {
        R.pair_a[rb + (np & ~(kWave - 1)) + lane] = my_a;
        R.pair_b[rb + (np & ~(kWave - 1)) + lane] = my_b;
    }
    // the lowest unused unit, if any: indices <= hw are used, above it the unmarked ones
    const int32_t left_k = 2 * np < nu ? next_free(hw, -1) : nu;
    if (lane == 0) {
        const int32_t left = left_k < nu ? left_k : -1;
        int32_t left_nnz = 0;
        if (left >= 0) {
            left_nnz = run_count(R, R.rest[rb + 2 * left]);
            if (2 * left + 1 < nrest) left_nnz += run_count(R, R.rest[rb + 2 * left + 1]);
        }
        const bool keep_left = left >= 0 && left_nnz >= unit_min_nnz;
        O.n_tcsub[w] = kept_total;
        O.n_tcval[w] = tcval;
        O.n_pairs[w] = np;
        O.n_groups[w] = np + (keep_left ? 1 : 0);
        O.leftover[w] = left;
        O.fallback[w] = (left >= 0 && !keep_left) ? 1 : 0;
        O.n_rest[w] = nrest;
    }
}

// ---------------------------------------------------------------- 3b. large windows, one workgroup each
//
// The decisions of k_window_decide with NW waves per window, for the windows
// with thousands of columns (R-MAT's hub windows: up to ~49 000 runs and
// ~23 000 units, 10.5 ms on one wave):
//   a-e  counts, the (count desc, col asc) order, TC chunks, the non-TC list and
//        the units run over the whole workgroup (stable counting sorts and
//        compactions with per-wave segments and workgroup prefix sums);
//   f    the libstdc++ introsort of the units: its top log2(NW) levels of
//        partitions run as a tree (level l on 2^l waves), then each wave sorts
//        one of the NW ranges alone; ranges are disjoint, so the result is the
//        sequential one.  The final insertion pass is a workgroup stable
//        counting sort by key;
//   g    the pairing on one wave, in scalar registers, with the active-unit
//        flags as a bitmap: the 64-flag word that holds the high-water mark
//        stays in two SGPRs, so a partner search is a masked find-first-zero
//        and a word load only when the mark passes the word.  Windows of more
//        than kHostPairUnits units stop before it: route_core pairs them on
//        the host (pair_on_host) and k_window_finish completes them.

constexpr int kKeyBuckets = 64;      // unit keys (nnz of two runs of a <= 16-row window) are < 33

// exclusive prefix sum of v over the workgroup (thread order); `red` holds
// NW + 1 ints; every thread gets the block total
template <int NW>
__device__ inline int32_t block_excl_scan(int32_t v, int32_t* red, int32_t& total) {
    const int lane = threadIdx.x & (kWave - 1), wid = threadIdx.x / kWave;
    int32_t x = v;
    for (int off = 1; off < kWave; off <<= 1) {
        const int32_t y = __shfl_up(x, off);
        if (lane >= off) x += y;
    }
    if (lane == kWave - 1) red[wid] = x;
    __syncthreads();
    if (threadIdx.x == 0) {
        int32_t acc = 0;
        for (int w = 0; w < NW; ++w) {
            const int32_t c = red[w];
            red[w] = acc;
            acc += c;
        }
        red[NW] = acc;
    }
    __syncthreads();
    const int32_t r = red[wid] + x - v;
    total = red[NW];
    __syncthreads();
    return r;
}

// one wave: libstdc++'s introsort loop on [first, last) with depth budget `depth`
// (lsort::sort's steps, the partitions in closed form), without the final pass
__device__ void wave_introsort_range(uint32_t* a, int32_t first, int32_t last, int depth, int32_t* __restrict__ tmp,
                                     int32_t* stack, PartLds& pl) {
    const WaveSync sync{};
    const int lane = threadIdx.x & (kWave - 1);
    if (lane == 0) {
        stack[0] = first;
        stack[1] = last;
        stack[2] = depth;
    }
    sync();
    int sp = 1;
    while (sp > 0) {
        --sp;
        int32_t f = stack[3 * sp], l = stack[3 * sp + 1], d = stack[3 * sp + 2];
        sync();
        while (l - f > lsort::kThreshold) {
            if (d == 0) {
                if (lane == 0) lsort::heap_sort(a + f, l - f, lsort::KeyDesc{});
                sync();
                break;
            }
            --d;
            const int32_t cut = wave_partition<WaveSync>(a, f, l, tmp + f, pl);
            if (lane == 0) {
                stack[3 * sp] = cut;
                stack[3 * sp + 1] = l;
                stack[3 * sp + 2] = d;
            }
            sync();
            ++sp;
            l = cut;
        }
    }
}

// bitmap of active units: LDS (plain loads, in wave order) or global memory
// (atomic loads: the marks are atomic ORs at the L2)
struct UnitFlags {
    uint64_t* w;
    bool global;
    __device__ uint64_t load(int32_t i) const {
        return global ? __hip_atomic_load(w + i, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) : w[i];
    }
    __device__ void mark(int32_t k) const {
        if (global) atomicOr((unsigned long long*)(w + (k >> 6)), 1ull << (k & 63));
        else atomicOr((unsigned long long*)(w + (k >> 6)), 1ull << (k & 63));
    }
};

template <int NW>
__global__ __launch_bounds__(64 * NW) void k_window_decide_wg(RunArrays R, WinOut O, int64_t W, int window, int density,
                                                             float tc_min_util, int skip_tc, int unit_min_nnz,
                                                             const int32_t* __restrict__ wlist, int lds_cap,
                                                             int32_t host_units) {
    constexpr int T = kWave * NW;
    extern __shared__ __align__(16) uint32_t dyn_wg[];
    __shared__ int32_t hist[17], base[17];
    __shared__ int32_t wcnt[NW][17];                 // per wave and run count: runs, then write offsets
    __shared__ int32_t kcnt[NW][kKeyBuckets];        // per wave and unit key: units, then write offsets
    __shared__ int32_t red[NW + 1];
    __shared__ int32_t s_ncand;
    __shared__ int32_t sort_stack[NW][3 * 66];
    __shared__ PartLds part[NW];
    __shared__ int32_t rng[NW][3], rng2[NW][3];      // the sort's ranges after each tree level
    const int64_t w = wlist[blockIdx.x];
    const int tid = threadIdx.x, lane = tid & (kWave - 1), wid = tid / kWave;
    const uint64_t lt = lanemask_lt();
    const int32_t rb = R.win_rb[w], re = R.win_rb[w + 1];
    const int32_t nc = re - rb;

    // a. histogram of run counts (1..16), per wave
    for (int i = tid; i < 17; i += T) hist[i] = 0;
    for (int i = tid; i < NW * 17; i += T) wcnt[i / 17][i % 17] = 0;
    __syncthreads();
    const int32_t seg = (nc + NW - 1) / NW;
    const int32_t s0 = rb + wid * seg, s1 = min(re, s0 + seg);
    for (int32_t i = s0 + lane; i < s1; i += kWave) atomicAdd(&wcnt[wid][run_count(R, i)], 1);
    __syncthreads();
    if (tid == 0) {
        for (int c = 1; c <= 16; ++c) {
            int32_t t = 0;
            for (int v = 0; v < NW; ++v) t += wcnt[v][c];
            hist[c] = t;
        }
        int32_t acc = 0;
        for (int c = 16; c >= 1; --c) {
            base[c] = acc;
            acc += hist[c];
        }
        int32_t ncand = 0;
        if (!skip_tc)
            for (int c = 16; c >= 1 && c >= density; --c) ncand += hist[c];
        s_ncand = ncand;
    }
    __syncthreads();
    // b. stable counting sort by count desc: wave v's runs of count c go after
    // every wave's runs of larger counts and after waves < v's runs of count c
    if (tid >= 1 && tid <= 16) {
        int32_t off = base[tid];
        for (int v = 0; v < NW; ++v) {
            const int32_t c = wcnt[v][tid];
            wcnt[v][tid] = off;
            off += c;
        }
    }
    __syncthreads();
    for (int32_t i0 = s0; i0 < s1; i0 += kWave) {
        const int32_t i = i0 + lane;
        const bool on = i < s1;
        const int c = on ? run_count(R, i) : 0;
        uint64_t rem = ballot(on);
        while (rem) {
            const int leader = __ffsll((long long)rem) - 1;
            const int cc = __shfl(c, leader);
            const uint64_t grp = ballot(on && c == cc);
            const int32_t at = wcnt[wid][cc];
            if (on && c == cc) R.order[rb + at + __popcll(grp & lt)] = i;
            if (lane == 0) wcnt[wid][cc] = at + __popcll(grp);
            rem &= ~grp;
        }
    }
    __syncthreads();

    // c. TC chunks of 8 over the candidates (block_online.cpp:215-253)
    const int32_t ncand = s_ncand;
    const int32_t nchunk = (ncand + 7) / 8;
    int32_t kept_total = 0, tcval = 0;
    for (int32_t j0 = 0; j0 < nchunk; j0 += T) {
        const int32_t j = j0 + tid;
        bool keep = false;
        int32_t nnz = 0;
        if (j < nchunk) {
            const int32_t e = (8 * j + 8 < ncand) ? 8 * j + 8 : ncand;
            for (int32_t k = 8 * j; k < e; ++k) nnz += run_count(R, R.order[rb + k]);
            const float util = (float)nnz / (float)(window * 8);
            keep = !(util < tc_min_util);
            for (int32_t k = 8 * j; k < e; ++k) R.in_tc[R.order[rb + k]] = keep ? 1 : 0;
        }
        int32_t nk = 0, sv = 0;
        const int32_t pos = block_excl_scan<NW>(keep ? 1 : 0, red, nk);
        if (keep) R.tc_chunk[rb + kept_total + pos] = j;
        block_excl_scan<NW>(keep ? nnz : 0, red, sv);
        kept_total += nk;
        tcval += sv;
    }
    for (int32_t k = ncand + tid; k < nc; k += T) R.in_tc[R.order[rb + k]] = 0;
    __syncthreads();

    // d. non-TC runs in (count desc, col asc) order
    int32_t nrest = 0;
    for (int32_t k0 = 0; k0 < nc; k0 += T) {
        const int32_t k = k0 + tid;
        const int32_t run = k < nc ? R.order[rb + k] : 0;
        const bool on = k < nc && !R.in_tc[run];
        int32_t n_on = 0;
        const int32_t pos = block_excl_scan<NW>(on ? 1 : 0, red, n_on);
        if (on) R.rest[rb + nrest + pos] = run;
        nrest += n_on;
    }
    __syncthreads();

    // e. units and their order key; the active-unit bitmap cleared
    const int32_t nu = (nrest + 1) / 2;
    const bool in_lds = lds_cap > 0 && nu <= lds_cap;
    uint32_t* units = in_lds ? dyn_wg : R.units_g + rb;
    for (int32_t k = tid; k < nu; k += T) {
        int32_t nnz = run_count(R, R.rest[rb + 2 * k]);
        if (2 * k + 1 < nrest) nnz += run_count(R, R.rest[rb + 2 * k + 1]);
        units[k] = ((uint32_t)nnz << 24) | (uint32_t)k;
    }
    // bitmap: in LDS after the keys (8-byte aligned: lds_cap is even), else in
    // used_g's bytes of this window rounded up to 8 bytes (nc >= nu / 8 + 8)
    UnitFlags fl;
    fl.global = !in_lds;
    fl.w = in_lds ? (uint64_t*)(dyn_wg + lds_cap)
                  : (uint64_t*)(((uintptr_t)(R.used_g + rb) + 7) & ~(uintptr_t)7);
    const int32_t nwords = (nu + 63) / 64;
    for (int32_t k = tid; k < nwords; k += T) fl.w[k] = 0;
    __syncthreads();

    // f. libstdc++ order of the units
    uint32_t* a = units;
    int32_t* tmp = R.pair_a + rb;                    // nc >= nu entries of scratch
    if (nu <= lsort::kThreshold) {
        if (tid == 0) lsort::sort(a, nu, lsort::KeyDesc{}, sort_stack[0]);
        __syncthreads();
    } else {
        if (tid == 0) {
            rng[0][0] = 0;
            rng[0][1] = nu;
            rng[0][2] = 2 * lsort::floor_log2(nu);
            for (int v = 1; v < NW; ++v) rng[v][0] = rng[v][1] = rng[v][2] = 0;
        }
        __syncthreads();
        // tree levels: wave v < 2^l partitions range v into ranges 2v and 2v + 1
        for (int nr = 1; 2 * nr <= NW; nr *= 2) {
            if (wid < nr) {
                const int32_t f = rng[wid][0], l = rng[wid][1], d = rng[wid][2];
                int32_t c0[3] = {f, l, d}, c1[3] = {0, 0, 0};
                if (l - f > lsort::kThreshold && d > 0) {
                    const int32_t cut = wave_partition<WaveSync>(a, f, l, tmp + f, part[wid]);
                    c0[1] = cut;
                    c0[2] = d - 1;
                    c1[0] = cut;
                    c1[1] = l;
                    c1[2] = d - 1;
                }
                if (lane == 0)
                    for (int q = 0; q < 3; ++q) {
                        rng2[2 * wid][q] = c0[q];
                        rng2[2 * wid + 1][q] = c1[q];
                    }
            }
            __syncthreads();
            if (tid < 2 * nr * 3) rng[tid / 3][tid % 3] = rng2[tid / 3][tid % 3];
            __syncthreads();
        }
        wave_introsort_range(a, rng[wid][0], rng[wid][1], rng[wid][2], tmp, sort_stack[wid], part[wid]);
        __syncthreads();
        // the final insertion pass = a stable sort by key (descending), per-wave
        // segments in order
        for (int i = tid; i < NW * kKeyBuckets; i += T) kcnt[i / kKeyBuckets][i % kKeyBuckets] = 0;
        __syncthreads();
        const int32_t useg = (nu + NW - 1) / NW;
        const int32_t u0 = wid * useg, u1 = min(nu, u0 + useg);
        for (int32_t p = u0 + lane; p < u1; p += kWave) atomicAdd(&kcnt[wid][a[p] >> 24], 1);
        __syncthreads();
        if (tid == 0) {
            int32_t acc = 0;
            for (int k = kKeyBuckets - 1; k >= 0; --k)
                for (int v = 0; v < NW; ++v) {
                    const int32_t c = kcnt[v][k];
                    kcnt[v][k] = acc;
                    acc += c;
                }
        }
        __syncthreads();
        uint32_t* out = (uint32_t*)tmp;
        for (int32_t b0 = u0; b0 < u1; b0 += kWave) {
            const int32_t p = b0 + lane;
            const bool on = p < u1;
            const uint32_t v = on ? a[p] : 0u;
            const uint32_t k = v >> 24;
            uint64_t rem = ballot(on);
            while (rem) {
                const int leader = __ffsll((long long)rem) - 1;
                const uint32_t kk = (uint32_t)__shfl((int)k, leader);
                const uint64_t grp = ballot(on && k == kk);
                const int32_t at = kcnt[wid][kk];
                if (on && k == kk) out[at + __popcll(grp & lt)] = v;
                if (lane == 0) kcnt[wid][kk] = at + __popcll(grp);
                rem &= ~grp;
            }
        }
        __syncthreads();
        for (int32_t p = tid; p < nu; p += T) a[p] = out[p];
        __syncthreads();
    }

    // g. pairing (block_online.cpp:365-391) on wave 0: in sorted order, a unit
    // not yet taken (index > hw) takes the first index above hw that no earlier
    // active unit holds, other than itself.  hw only grows, so the search keeps
    // the 64-flag word that holds hw + 1 in scalar registers (indices sb ..
    // sb + 63): a unit marked inside that word sets its bit there only (the
    // word is never read again once hw passes it), one marked beyond it sets
    // its bit in the bitmap, from which the word is loaded when hw reaches it.
    // A ballot per 64 units skips those already at or below hw.  Pairs are
    // collected in two registers (pair k in lane k % 64) and stored 64 at a time.
    if (host_units > 0 && nu > host_units) {
        // the pairing runs on the host (route_core, pair_on_host): the sorted
        // keys go to units_g, the counts known so far to O
        if (in_lds)
            for (int32_t p = tid; p < nu; p += T) R.units_g[rb + p] = a[p];
        if (tid == 0) {
            O.n_tcsub[w] = kept_total;
            O.n_tcval[w] = tcval;
            O.n_rest[w] = nrest;
        }
        return;
    }
    if (wid != 0) return;
    int32_t hw = -1, np = 0, left_k = nu;
    auto pairing = [&](auto bits) {
        auto load_word = [&](int32_t wi) -> uint64_t {   // word wi, indices >= nu read as taken
            const int32_t i0 = 64 * wi;
            if (i0 >= nu) return ~0ull;
            uint64_t u = bits.load(wi);
            if (nu - i0 < 64) u |= ~0ull << (nu - i0);
            return u;
        };
        int32_t sb = 0;
        uint64_t sw = load_word(0);
        // first index > hw whose flag is clear, other than `except` (nu if none);
        // leaves sb / sw on the word searched last
        auto search = [&](int32_t except) -> int32_t {
            uint64_t m = ~sw & (~0ull << (hw + 1 - sb));
            const uint32_t de = (uint32_t)(except - sb);
            if (de < 64u) m &= ~(1ull << de);
            while (m == 0) {
                sb += 64;
                if (sb >= nu) return nu;
                sw = load_word(sb >> 6);
                m = ~sw;
                const uint32_t d2 = (uint32_t)(except - sb);
                if (d2 < 64u) m &= ~(1ull << d2);
            }
            return sb + (int32_t)__builtin_ctzll(m);
        };
        uint32_t va = 0, vb = 0;
        int32_t last = -1;                           // a unit that found no partner (then no later one can)
        for (int32_t i0 = 0; i0 < nu && last < 0; i0 += kWave) {
            const int32_t mine = i0 + lane < nu ? (int32_t)(a[i0 + lane] & 0xffffffu) : -1;
            uint64_t cand = __ballot(mine > hw);
            while (cand) {
                const int j = (int)__builtin_ctzll(cand);
                cand &= cand - 1;
                const int32_t v = __builtin_amdgcn_readlane(mine, j);
                if (v <= hw) continue;                   // already a partner
                const int32_t p = search(v);
                if (p >= nu) {                           // every index above hw but v is taken
                    last = v;
                    break;
                }
                const uint32_t dv = (uint32_t)(v - sb);
                if (dv < 64u) sw |= 1ull << dv;
                else bits.mark(v, lane);
                if (lane == (np & (kWave - 1))) {
                    va = (uint32_t)v;
                    vb = (uint32_t)p;
                }
                hw = p;
                if (((hw + 1) >> 6) != (sb >> 6)) {      // keep the word of hw + 1 at hand
                    sb = (hw + 1) & ~63;
                    sw = load_word(sb >> 6);
                }
                if ((++np & (kWave - 1)) == 0) {
                    R.pair_a[rb + np - kWave + lane] = (int32_t)va;
                    R.pair_b[rb + np - kWave + lane] = (int32_t)vb;
                }
            }
        }
        if (lane < (np & (kWave - 1))) {
            R.pair_a[rb + (np & ~(kWave - 1)) + lane] = (int32_t)va;
            R.pair_b[rb + (np & ~(kWave - 1)) + lane] = (int32_t)vb;
        }
        if (last >= 0) left_k = last;            // the one index above hw still free
        else if (2 * np < nu) left_k = search(-1);
    };
    if (in_lds) {
        struct {
            uint64_t* w;
            __device__ uint64_t load(int32_t i) const {
                const uint64_t x = w[i];
                return (uint64_t)(uint32_t)__builtin_amdgcn_readfirstlane((int)(uint32_t)x) |
                       ((uint64_t)(uint32_t)__builtin_amdgcn_readfirstlane((int)(uint32_t)(x >> 32)) << 32);
            }
            __device__ void mark(int32_t k, int ln) const {
                if (ln == 0) atomicOr((unsigned long long*)(w + (k >> 6)), 1ull << (k & 63));
            }
        } bits{(uint64_t*)(dyn_wg + lds_cap)};
        pairing(bits);
    } else {
        struct {
            uint64_t* w;
            __device__ uint64_t load(int32_t i) const {
                const uint64_t x = __hip_atomic_load(w + i, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                return (uint64_t)(uint32_t)__builtin_amdgcn_readfirstlane((int)(uint32_t)x) |
                       ((uint64_t)(uint32_t)__builtin_amdgcn_readfirstlane((int)(uint32_t)(x >> 32)) << 32);
            }
            __device__ void mark(int32_t k, int ln) const {
                if (ln == 0) atomicOr((unsigned long long*)(w + (k >> 6)), 1ull << (k & 63));
            }
        } bits{fl.w};
        pairing(bits);
    }
    if (lane == 0) {
        const int32_t left = left_k < nu ? left_k : -1;
        int32_t left_nnz = 0;
        if (left >= 0) {
            left_nnz = run_count(R, R.rest[rb + 2 * left]);
            if (2 * left + 1 < nrest) left_nnz += run_count(R, R.rest[rb + 2 * left + 1]);
        }
        const bool keep_left = left >= 0 && left_nnz >= unit_min_nnz;
        O.n_tcsub[w] = kept_total;
        O.n_tcval[w] = tcval;
        O.n_pairs[w] = np;
        O.n_groups[w] = np + (keep_left ? 1 : 0);
        O.leftover[w] = left;
        O.fallback[w] = (left >= 0 && !keep_left) ? 1 : 0;
        O.n_rest[w] = nrest;
    }
}

// The pairing of a window decided with host_pairing, on the host: the loop of
// k_window_decide_wg's step g over the sorted unit keys (unit index in the low
// 24 bits).  In sorted order, a unit above the high-water mark hw takes the
// first index above hw that no earlier such unit holds, other than itself, and
// hw moves there; a unit that finds none stays over, as does, when every unit
// got a turn, the first free index above hw.  One sequential pass: on the
// device this loop ran on one wave at about 115 ns per unit (R-MAT scale 18's
// hub window, 23 355 units: 2.65 ms).
// (windows of more than kHostPairUnits units, the two largest size classes:
// their pairing takes up to milliseconds each on the device; on the host the
// loop takes about 1 ns per unit, and the windows' keys and pairs travel in
// one copy each way)
constexpr int32_t kHostPairUnits = 6144;

static void pair_on_host(const uint32_t* keys, int32_t nu, int32_t* pa, int32_t* pb, int32_t& np, int32_t& left_k) {
    std::vector<uint8_t> active((size_t)nu, 0);
    int32_t hw = -1, scan = 0, last = -1;
    np = 0;
    // first index > hw that is not active and is not `except` (nu if none)
    auto search = [&](int32_t except) {
        if (scan < hw + 1) scan = hw + 1;
        int32_t q = scan;
        while (q < nu && (active[q] || q == except)) ++q;
        return q;
    };
    for (int32_t i = 0; i < nu; ++i) {
        const int32_t v = (int32_t)(keys[i] & 0xffffffu);
        if (v <= hw) continue;                      // already a partner
        const int32_t p = search(v);
        if (p >= nu) {                              // every index above hw but v is taken
            last = v;
            break;
        }
        active[v] = 1;
        pa[np] = v;
        pb[np] = p;
        ++np;
        hw = p;
    }
    left_k = nu;
    if (last >= 0) left_k = last;
    else if (2 * np < nu) left_k = search(-1);
}

// the keys of the windows paired on the host, packed: block b copies the
// sorted unit keys of window tab[4 b] to keys[tab[4 b + 1] ..]
__global__ __launch_bounds__(256) void k_pack_units(RunArrays R, WinOut O, const int32_t* __restrict__ tab,
                                                    uint32_t* __restrict__ keys) {
    const int64_t w = tab[4 * blockIdx.x];
    const int32_t off = tab[4 * blockIdx.x + 1];
    const int32_t rb = R.win_rb[w], nu = (O.n_rest[w] + 1) / 2;
    for (int32_t k = threadIdx.x; k < nu; k += blockDim.x) keys[off + k] = R.units_g[rb + k];
}

// the rest of a window paired on the host: block b places the pairs of window
// tab[4 b] (tab[4 b + 2] of them, packed from tab[4 b + 1]) in pair_a / pair_b
// and writes its counts; tab[4 b + 3] is the leftover unit (nu if none)
__global__ __launch_bounds__(256) void k_window_finish(RunArrays R, WinOut O, const int32_t* __restrict__ tab,
                                                       const int32_t* __restrict__ pa, const int32_t* __restrict__ pb,
                                                       int unit_min_nnz) {
    const int64_t w = tab[4 * blockIdx.x];
    const int32_t off = tab[4 * blockIdx.x + 1], np = tab[4 * blockIdx.x + 2], left_k = tab[4 * blockIdx.x + 3];
    const int32_t rb = R.win_rb[w];
    for (int32_t k = threadIdx.x; k < np; k += blockDim.x) {
        R.pair_a[rb + k] = pa[off + k];
        R.pair_b[rb + k] = pb[off + k];
    }
    if (threadIdx.x != 0) return;
    const int32_t nrest = O.n_rest[w], nu = (nrest + 1) / 2;
    const int32_t left = left_k < nu ? left_k : -1;
    int32_t left_nnz = 0;
    if (left >= 0) {
        left_nnz = run_count(R, R.rest[rb + 2 * left]);
        if (2 * left + 1 < nrest) left_nnz += run_count(R, R.rest[rb + 2 * left + 1]);
    }
    const bool keep_left = left >= 0 && left_nnz >= unit_min_nnz;
    O.n_pairs[w] = np;
    O.n_groups[w] = np + (keep_left ? 1 : 0);
    O.leftover[w] = left;
    O.fallback[w] = (left >= 0 && !keep_left) ? 1 : 0;
}

__global__ void k_widen(const uint8_t* __restrict__ in, int64_t n, int32_t* __restrict__ out) {
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
        out[i] = in[i];
}

__global__ void k_part_counts(const int32_t* __restrict__ n_tcsub, const int32_t* __restrict__ n_groups, int64_t W,
                              int part_t, int32_t* __restrict__ n_tparts, int32_t* __restrict__ n_tiles) {
    const int64_t w = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (w >= W) return;
    const int32_t s = n_tcsub[w];
    n_tparts[w] = s == 0 ? 0 : (s <= part_t ? 1 : s / part_t + (s % part_t ? 1 : 0));
    n_tiles[w] = (n_groups[w] + 3) / 4;
}

// ---------------------------------------------------------------- 5. emission
struct Emit {
    // sorted edges
    const int32_t* sval;     // edge id at each sorted position
    const int32_t* erow;     // row of each edge
    const float* val;
    // window prefix sums
    const int32_t* tc_sub_off;
    const int32_t* tc_val_off;
    const int32_t* tc_part_off;
    const int32_t* grp_off;       // kept 2:4 groups per window (= s_offsets)
    const int32_t* tile_off;
    const int32_t* grp_all_off;   // every group the reference builds, per window
    const int32_t* gmap;          // group (all) -> index among kept groups
    uint8_t* keep;                // group (all) stays a 2:4 group (TUNED mode drops sparse ones)
    // outputs (block_online.cpp:844-868 order)
    int32_t *t_row_offset, *t_block_offset, *t_column;
    float* t_value;
    int32_t *t_window_row, *t_atomic;
    int64_t* t_binary;
    int32_t* s_column;
    uint8_t* s_mask;
    float* s_value;
    int8_t* s_pos;
    int32_t *s_window, *s_offsets;
    int32_t *s_tile_window, *s_tile_group, *s_tile_column, *s_packed_a, *s_packed_meta, *s_window_tile_offset;
    uint8_t* res_flag;       // per edge: goes to the residual route
    // optional (null = not requested): the CSR edge behind each emitted value,
    // t_edge [TV] beside t_value, s_edge [G, window, 2] beside s_value (-1 = empty)
    int32_t* t_edge = nullptr;
    int32_t* s_edge = nullptr;
};

// TC: one thread per (window, kept chunk) = sub-block
// values of each TC sub-block (the runs of its chunk), in sub-block order
__global__ void k_tc_counts(RunArrays R, Emit E, int64_t W, int64_t S, int32_t* __restrict__ cnt) {
    const int64_t s = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= S) return;
    const int64_t w = upper_bound_window(E.tc_sub_off, W, (int32_t)s);
    const int32_t rb = R.win_rb[w];
    const int32_t j = R.tc_chunk[rb + (int32_t)(s - E.tc_sub_off[w])];
    const int32_t nw = R.win_rb[w + 1] - rb;
    int32_t c = 0;
    for (int q = 0; q < 8; ++q) {   // chunk j = order[8j, 8j+8) cut at the first non-candidate
        const int32_t pos = 8 * j + q;
        if (pos >= nw) break;
        const int32_t run = R.order[rb + pos];
        if (!R.in_tc[run]) break;
        c += run_count(R, run);
    }
    cnt[s] = c;
}

__global__ void k_emit_tc(RunArrays R, WinOut O, Emit E, int64_t W, int window, int part_t, int64_t S) {
    const int64_t s = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= S) return;
    const int64_t w = upper_bound_window(E.tc_sub_off, W, (int32_t)s);
    const int32_t rb = R.win_rb[w];
    const int32_t k = (int32_t)(s - E.tc_sub_off[w]);
    const int32_t j = R.tc_chunk[rb + k];
    // the chunk's runs, sorted by column (<= 8)
    int32_t runs[8], cols[8];
    int n = 0;
    // chunk j = order[8j, 8j+8) cut at the first non-candidate (in_tc == 0)
    for (int q = 0; q < 8; ++q) {
        const int32_t pos = 8 * j + q;
        if (pos >= R.win_rb[w + 1] - rb) break;
        const int32_t run = R.order[rb + pos];
        if (!R.in_tc[run]) break;
        runs[n] = run;
        cols[n] = R.col[run];
        ++n;
    }
    for (int a = 1; a < n; ++a) {   // insertion sort by column
        const int32_t rr = runs[a], cc = cols[a];
        int b = a - 1;
        while (b >= 0 && cols[b] > cc) {
            runs[b + 1] = runs[b];
            cols[b + 1] = cols[b];
            --b;
        }
        runs[b + 1] = rr;
        cols[b + 1] = cc;
    }
    for (int q = 0; q < 8; ++q) E.t_column[8 * s + q] = q < n ? cols[q] : -1;
    // value offset of this sub-block: the prefix sum of k_tc_counts, which equals
    // the window's offset plus the values of its earlier kept chunks
    int32_t voff = E.t_block_offset[s];
    // bits and values in (row, slot) order
    int32_t ptr[8];
    for (int q = 0; q < n; ++q) ptr[q] = R.start[runs[q]];
    const int64_t row0 = w * window;
    uint64_t lo = 0, hi = 0;
    for (int lr = 0; lr < window; ++lr) {
        for (int q = 0; q < n; ++q) {
            if (ptr[q] < R.start[runs[q] + 1]) {
                const int32_t e = E.sval[ptr[q]];
                if (E.erow[e] - row0 == lr) {
                    if (lr >= 8) hi |= 1ull << ((lr - 8) * 8 + q);
                    else lo |= 1ull << (lr * 8 + q);
                    if (E.t_edge) E.t_edge[voff] = e;
                    E.t_value[voff++] = E.val[e];
                    ++ptr[q];
                }
            }
        }
    }
    E.t_binary[2 * s] = (int64_t)lo;
    E.t_binary[2 * s + 1] = (int64_t)hi;
    // parts: the thread of the first sub-block of each part writes it
    const int32_t nsub = O.n_tcsub[w];
    if (k % part_t == 0) {
        const int32_t p = E.tc_part_off[w] + k / part_t;
        E.t_row_offset[p] = (int32_t)s;
        E.t_window_row[p] = (int32_t)(w * 16);
        E.t_atomic[p] = nsub <= part_t ? O.fallback[w] : 1;
    }
}

// the runs (<= 4) of group ga (index among all groups), sorted by column
__device__ inline int group_runs(const RunArrays& R, const WinOut& O, const Emit& E, int64_t ga, int64_t W,
                                 int64_t& w, int32_t (&runs)[4], int32_t (&cols)[4]) {
    w = upper_bound_window(E.grp_all_off, W, (int32_t)ga);
    const int32_t rb = R.win_rb[w];
    const int32_t k = (int32_t)(ga - E.grp_all_off[w]);
    int32_t ua, ub = -1;
    if (k < O.n_pairs[w]) {
        ua = R.pair_a[rb + k];
        ub = R.pair_b[rb + k];
    } else {
        ua = O.leftover[w];
    }
    // columns of the unit(s): rest[2u], rest[2u+1] if that unit is a pair
    const int32_t nr = O.n_rest[w];
    int n = 0;
    const int32_t us[2] = {ua, ub};
    for (int t = 0; t < 2; ++t) {
        const int32_t u = us[t];
        if (u < 0) continue;
        for (int h = 0; h < 2; ++h) {
            const int32_t pos = 2 * u + h;
            if (pos >= nr) break;
            runs[n] = R.rest[rb + pos];
            cols[n] = R.col[runs[n]];
            ++n;
        }
    }
    for (int a = 1; a < n; ++a) {
        const int32_t rr = runs[a], cc = cols[a];
        int b = a - 1;
        while (b >= 0 && cols[b] > cc) {
            runs[b + 1] = runs[b];
            cols[b + 1] = cols[b];
            --b;
        }
        runs[b + 1] = rr;
        cols[b + 1] = cc;
    }
    return n;
}

// TUNED mode: a group whose kept slots (at most 2 hits per row) number fewer
// than min_nnz leaves the 2:4 route; min_nnz = 0 keeps every group.
__global__ void k_group_keep(RunArrays R, WinOut O, Emit E, int64_t W, int window, int min_nnz, int64_t GA) {
    const int64_t ga = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (ga >= GA) return;
    if (min_nnz <= 0) {
        E.keep[ga] = 1;
        return;
    }
    int64_t w;
    int32_t runs[4], cols[4];
    const int n = group_runs(R, O, E, ga, W, w, runs, cols);
    int hits[16] = {0};
    for (int q = 0; q < n; ++q)
        for (int32_t i = R.start[runs[q]]; i < R.start[runs[q] + 1]; ++i) {
            const int lr = (int)(E.erow[E.sval[i]] - w * window);
            if (lr >= 0 && lr < 16) ++hits[lr];
        }
    int kept = 0;
    for (int lr = 0; lr < 16; ++lr) kept += hits[lr] < 2 ? hits[lr] : 2;
    E.keep[ga] = kept >= min_nnz ? 1 : 0;
}

// kept-group offsets per window and 2:4 tiles per window
__global__ void k_group_offsets(const int32_t* __restrict__ grp_all_off, const int32_t* __restrict__ gmap, int64_t W,
                                int32_t* __restrict__ s_offsets, int32_t* __restrict__ n_tiles) {
    const int64_t w = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (w > W) return;
    s_offsets[w] = gmap[grp_all_off[w]];
    if (w < W) n_tiles[w] = (gmap[grp_all_off[w + 1]] - gmap[grp_all_off[w]] + 3) / 4;
}

// 2:4 groups: one thread per group the reference builds; kept groups are
// emitted at their kept index, dropped ones send every hit to the residual route
__global__ void k_emit_groups(RunArrays R, WinOut O, Emit E, int64_t W, int window, int exact, int64_t GA) {
    const int64_t ga = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (ga >= GA) return;
    int64_t w;
    int32_t runs[4], cols[4];
    const int n = group_runs(R, O, E, ga, W, w, runs, cols);
    if (!E.keep[ga]) {
        for (int q = 0; q < n; ++q)
            for (int32_t i = R.start[runs[q]]; i < R.start[runs[q] + 1]; ++i) E.res_flag[E.sval[i]] = 1;
        return;
    }
    const int64_t g = E.gmap[ga];
    for (int q = 0; q < 4; ++q) E.s_column[4 * g + q] = q < n ? cols[q] : -1;
    E.s_window[g] = (int32_t)w;
    int32_t ptr[4];
    for (int q = 0; q < n; ++q) ptr[q] = R.start[runs[q]];
    const int64_t row0 = w * window;
    for (int lr = 0; lr < window; ++lr) {
        uint8_t mask = 0;
        int used = 0;
        float v[2] = {0.f, 0.f};
        int8_t ps[2] = {-1, -1};
        int32_t ed[2] = {-1, -1};
        for (int q = 0; q < n; ++q) {
            if (ptr[q] >= R.start[runs[q] + 1]) continue;
            const int32_t e = E.sval[ptr[q]];
            if (E.erow[e] - row0 != lr) continue;
            ++ptr[q];
            mask |= (uint8_t)(1u << q);
            if (used < 2) {
                v[used] = E.val[e];
                ed[used] = e;
                ps[used] = (int8_t)q;
            } else if (exact) {
                E.res_flag[e] = 1;
            }
            ++used;
        }
        const int64_t b = g * window + lr;
        E.s_mask[b] = mask;
        E.s_value[2 * b] = v[0];
        E.s_value[2 * b + 1] = v[1];
        E.s_pos[2 * b] = ps[0];
        E.s_pos[2 * b + 1] = ps[1];
        if (E.s_edge) {
            E.s_edge[2 * b] = ed[0];
            E.s_edge[2 * b + 1] = ed[1];
        }
    }
}

// columns in neither set: the dropped leftover unit
__global__ void k_flag_dropped(RunArrays R, WinOut O, Emit E, int64_t W) {
    const int64_t w = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (w >= W || !O.fallback[w]) return;
    const int32_t rb = R.win_rb[w];
    const int32_t u = O.leftover[w];
    const int32_t nr = O.n_rest[w];
    for (int h = 0; h < 2; ++h) {
        const int32_t pos = 2 * u + h;
        if (pos >= nr) break;
        const int32_t run = R.rest[rb + pos];
        for (int32_t i = R.start[run]; i < R.start[run + 1]; ++i) E.res_flag[E.sval[i]] = 1;
    }
}

// residual: per row counts of parts (long/short) and entries (block_online.cpp:415-439)
struct ResCounts {
    int32_t *nl, *ns, *el, *es;   // per row: long parts, short parts, long entries, short entries
};

__global__ __launch_bounds__(256) void k_res_count(const int32_t* __restrict__ rowptr, const uint8_t* __restrict__ flag,
                                                   int64_t rows, int part_c, int short_c, ResCounts RC) {
    const int64_t r = wave_row();
    const int lane = threadIdx.x & 63;
    if (r >= rows) return;
    int32_t n = 0;
    for (int32_t e = rowptr[r] + lane; e < rowptr[r + 1]; e += kWave) n += flag[e];
    for (int off = 32; off > 0; off >>= 1) n += __shfl_xor(n, off);
    if (lane != 0) return;
    int32_t nl = 0, ns = 0, el = 0, es = 0;
    if (n > 0) {
        if (n <= short_c) { ns = 1; es = n; }
        else if (n <= part_c) { nl = 1; el = n; }
        else {
            nl = n / part_c;
            el = nl * part_c;
            const int32_t rem = n % part_c;
            if (rem) {
                if (rem <= short_c) { ns = 1; es = rem; }
                else { nl += 1; el += rem; }
            }
        }
    }
    RC.nl[r] = nl;
    RC.ns[r] = ns;
    RC.el[r] = el;
    RC.es[r] = es;
}

struct ResOut {
    const int32_t *pl, *ps, *ql, *qs;     // prefix: long parts, short parts, long entries, short entries
    int32_t *c_row_offset, *c_row, *c_atomic, *c_column;
    float* c_value;
    int32_t *cs_row_offset, *cs_row, *cs_atomic, *cs_column;
    float* cs_value;
    int32_t* c_edge = nullptr;     // optional: the CSR edge of each long / short entry
    int32_t* cs_edge = nullptr;
};

__global__ __launch_bounds__(256) void k_res_emit(const int32_t* __restrict__ rowptr, const int32_t* __restrict__ col,
                                                  const float* __restrict__ val, const uint8_t* __restrict__ flag,
                                                  const int32_t* __restrict__ n_tcsub, int64_t rows, int window,
                                                  int part_c, int short_c, ResOut O) {
    const int64_t r = wave_row();
    const int lane = threadIdx.x & 63;
    if (r >= rows) return;
    const int32_t nl = O.pl[r + 1] - O.pl[r], ns = O.ps[r + 1] - O.ps[r];
    if (nl == 0 && ns == 0) return;
    const int32_t el = O.ql[r + 1] - O.ql[r];
    const int32_t base_atomic = n_tcsub[r / window] > 0 ? 1 : 0;
    const int32_t n = el + (O.qs[r + 1] - O.qs[r]);
    const bool split = n > part_c;
    // long parts
    for (int32_t j = lane; j < nl; j += kWave) {
        const int32_t p = O.pl[r] + j;
        O.c_row_offset[p] = O.ql[r] + j * part_c;
        O.c_row[p] = (int32_t)r;
        O.c_atomic[p] = split ? 1 : base_atomic;
    }
    if (ns && lane == 0) {
        const int32_t p = O.ps[r];
        O.cs_row_offset[p] = O.qs[r];
        O.cs_row[p] = (int32_t)r;
        O.cs_atomic[p] = split ? 1 : base_atomic;
    }
    // the row's flagged entries in CSR order: the first el to the long parts, the rest short
    const uint64_t lt = lanemask_lt();
    int32_t k0 = 0;
    for (int32_t b = rowptr[r]; b < rowptr[r + 1]; b += kWave) {
        const int32_t e = b + lane;
        const bool on = e < rowptr[r + 1] && flag[e];
        const uint64_t m = ballot(on);
        if (on) {
            const int32_t k = k0 + __popcll(m & lt);
            if (k < el) {
                O.c_column[O.ql[r] + k] = col[e];
                O.c_value[O.ql[r] + k] = val[e];
                if (O.c_edge) O.c_edge[O.ql[r] + k] = e;
            } else {
                O.cs_column[O.qs[r] + k - el] = col[e];
                O.cs_value[O.qs[r] + k - el] = val[e];
                if (O.cs_edge) O.cs_edge[O.qs[r] + k - el] = e;
            }
        }
        k0 += __popcll(m);
    }
}

__global__ void k_res_tail(ResOut O, int64_t rows, int64_t PL, int64_t PS) {
    O.c_row_offset[PL] = O.ql[rows];
    O.cs_row_offset[PS] = O.qs[rows];
}

__global__ void k_tail(Emit E, int64_t S, int64_t PT, int64_t TV) {
    E.t_row_offset[PT] = (int32_t)S;
    E.t_block_offset[S] = (int32_t)TV;
}

// mma.sp tile packing (block_online.cpp:796-841): one thread per (tile, row group)
__global__ void k_pack_tiles(Emit E, int64_t W, int window, int64_t T) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= T * 8) return;
    const int64_t t = i >> 3;
    const int rg = (int)(i & 7);
    const int64_t w = upper_bound_window(E.tile_off, W, (int32_t)t);
    const int32_t ge = E.grp_off[w + 1];
    const int32_t gb = E.grp_off[w] + 4 * (int32_t)(t - E.tile_off[w]);
    if (rg == 0) {
        E.s_tile_window[t] = (int32_t)w;
        E.s_tile_group[t] = gb;
        for (int k = 0; k < 16; ++k) {
            const int32_t g = gb + (k >> 2);
            E.s_tile_column[t * 16 + k] = g < ge ? E.s_column[4 * g + (k & 3)] : -1;
        }
    }
    uint32_t meta = 0;
    for (int ch = 0; ch < 4; ++ch) {
        for (int hf = 0; hf < 2; ++hf) {
            const int row = rg + 8 * hf;
            const int32_t g = gb + ch;
            float v0 = 0.f, v1 = 0.f;
            int p0 = 0, p1 = 1;
            if (g < ge && row < window) {
                const int64_t b = ((int64_t)g * window + row) * 2;
                const int q0 = E.s_pos[b], q1 = E.s_pos[b + 1];
                const float x0 = E.s_value[b], x1 = E.s_value[b + 1];
                if (q0 >= 0 && q1 >= 0) {
                    if (q0 <= q1) { p0 = q0; p1 = q1; v0 = x0; v1 = x1; }
                    else          { p0 = q1; p1 = q0; v0 = x1; v1 = x0; }
                } else if (q0 >= 0 || q1 >= 0) {
                    const int q = q0 >= 0 ? q0 : q1;
                    const float x = q0 >= 0 ? x0 : x1;
                    const int dummy = q == 0 ? 1 : 0;
                    if (dummy < q) { p0 = dummy; p1 = q; v1 = x; }
                    else           { p0 = q; p1 = dummy; v0 = x; }
                }
            }
            meta |= ((uint32_t)(p0 & 3) | ((uint32_t)(p1 & 3) << 2)) << (16 * hf + 4 * ch);
            const uint32_t packed = (uint32_t)f32_to_f16_rne(v0) | ((uint32_t)f32_to_f16_rne(v1) << 16);
            E.s_packed_a[((t * 8 + rg) * 4 + ch) * 2 + hf] = (int32_t)packed;
        }
    }
    E.s_packed_meta[t * 8 + rg] = (int32_t)meta;
}

inline unsigned nblk(int64_t n, int t = 256) {
    int64_t b = (n + t - 1) / t;
    return (unsigned)(b < 1 ? 1 : (b > (int64_t)1 << 30 ? (int64_t)1 << 30 : b));
}
inline unsigned nblk_gs(int64_t n, int t = 256) {   // grid-stride launches
    int64_t b = (n + t - 1) / t;
    if (b < 1) b = 1;
    if (b > 65536) b = 65536;
    return (unsigned)b;
}

int bit_width(int64_t x) {
    int b = 0;
    while (x > 0) {
        ++b;
        x >>= 1;
    }
    return b < 1 ? 1 : b;
}

// three streams per device (default priority, created once) on which the
// smaller size classes of the window decisions run beside the largest one
constexpr int kSide = 3;
int side_streams(hipStream_t (&out)[kSide]) {
    static std::mutex mu;
    static std::map<int, std::array<hipStream_t, kSide>> made;
    int dev = 0;
    TRSPMM_HIP(hipGetDevice(&dev));
    std::lock_guard<std::mutex> g(mu);
    auto it = made.find(dev);
    if (it == made.end()) {
        std::array<hipStream_t, kSide> st{};
        for (int i = 0; i < kSide; ++i)
            if (hipStreamCreateWithFlags(&st[i], hipStreamNonBlocking) != hipSuccess) {
                (void)hipGetLastError();
                for (int j = 0; j < i; ++j) (void)hipStreamDestroy(st[j]);
                set_error("hipStreamCreateWithFlags failed");
                return TRSPMM_EHIP;
            }
        it = made.emplace(dev, st).first;
    }
    for (int i = 0; i < kSide; ++i) out[i] = it->second[i];
    return TRSPMM_OK;
}

// events of one fork/join, destroyed with the scope (HIP releases a recorded
// event once its work is done); a scope left before the join (an error) waits
// for the side streams, so that no scratch is handed back under their work
struct EventSet {
    hipEvent_t e[kSide + 1] = {};
    hipStream_t side[kSide] = {};
    bool joined = false;
    int make() {
        for (auto& x : e) TRSPMM_HIP(hipEventCreateWithFlags(&x, hipEventDisableTiming));
        return TRSPMM_OK;
    }
    ~EventSet() {
        if (!joined)
            for (auto& x : side)
                if (x) (void)hipStreamSynchronize(x);
        for (auto& x : e)
            if (x) (void)hipEventDestroy(x);
    }
};

// ---------------------------------------------------------------- row reordering
// TRSPMM_ROUTE_REORDER: the rows in order of (first column, nonzeros
// descending, row), so that rows whose lowest column is the same hub sit in
// the same windows, the longest of them first; empty rows go last.  On a
// power-law graph a window then holds columns that reach the TC density or
// fill 2:4 groups, which one gathered B row serves for several rows.
__global__ void k_row_keys(const int32_t* __restrict__ rowptr, const int32_t* __restrict__ col, int64_t rows,
                           uint64_t* __restrict__ keys, int32_t* __restrict__ idx) {
    for (int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; r < rows; r += (int64_t)gridDim.x * blockDim.x) {
        const int32_t b = rowptr[r], e = rowptr[r + 1];
        const uint32_t first = e > b ? (uint32_t)col[b] : 0xffffffffu;
        keys[r] = ((uint64_t)first << 32) | (uint64_t)(0xffffffffu - (uint32_t)(e - b));
        idx[r] = (int32_t)r;
    }
}

__global__ void k_perm_counts(const int32_t* __restrict__ rowptr, const int32_t* __restrict__ order, int64_t rows,
                              int32_t* __restrict__ cnt) {
    for (int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; r < rows; r += (int64_t)gridDim.x * blockDim.x) {
        const int32_t src = order[r];
        cnt[r] = rowptr[src + 1] - rowptr[src];
    }
}

// one wave per routed row: the entries of its input row, and the input edge of each
__global__ __launch_bounds__(256) void k_perm_rows(const int32_t* __restrict__ rowptr, const int32_t* __restrict__ col,
                                                   const float* __restrict__ val, const int32_t* __restrict__ order,
                                                   const int32_t* __restrict__ prowptr, int64_t rows,
                                                   int32_t* __restrict__ pcol, float* __restrict__ pval,
                                                   int32_t* __restrict__ pedge) {
    const int64_t r = wave_row();
    if (r >= rows) return;
    const int lane = threadIdx.x & 63;
    const int32_t src = order[r], b = rowptr[src], n = rowptr[src + 1] - b, d = prowptr[r];
    for (int32_t i = lane; i < n; i += kWave) {
        pcol[d + i] = col[b + i];
        pval[d + i] = val[b + i];
        pedge[d + i] = b + i;
    }
}

// edge maps of the reordered routing: routed edge -> input edge (-1 stays)
__global__ void k_map_edges(int32_t* __restrict__ e, int64_t n, const int32_t* __restrict__ pedge) {
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
        const int32_t x = e[i];
        if (x >= 0) e[i] = pedge[x];
    }
}

// entries >= 0 of a[n] (column lists: -1 marks an empty slot), added to *out
__global__ __launch_bounds__(256) void k_count_valid(const int32_t* __restrict__ a, int64_t n,
                                                     unsigned long long* __restrict__ out) {
    unsigned long long c = 0;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
        c += a[i] >= 0;
    for (int off = 32; off > 0; off >>= 1) c += __shfl_xor(c, off);
    if ((threadIdx.x & 63) == 0 && c) atomicAdd(out, c);
}

}  // namespace

// ---------------------------------------------------------------- host driver
static int route_core(const int32_t* rowptr, const int32_t* col, const float* val, int64_t rows, int64_t nnz,
                      int64_t ncols, const trspmm_route_cfg& cfg, hipStream_t s, Meta& M) {
    const auto t_begin = std::chrono::steady_clock::now();
    StreamArena A(s);      // scratch: declared first, so that every buffer below is back before it waits
    TRSPMM_CHECK(cfg.window > 0 && cfg.window <= 16, "window must be in [1, 16]");
    TRSPMM_CHECK(cfg.wide > 0 && cfg.wide <= 32, "wide must be in [1, 32]");
    TRSPMM_CHECK(cfg.part_t > 0 && cfg.part_c > 0, "partSize_t and partSize_c must be positive");
    TRSPMM_CHECK(cfg.mode == TRSPMM_MODE_COMPAT || cfg.mode == TRSPMM_MODE_EXACT || cfg.mode == TRSPMM_MODE_TUNED,
                 "router mode must be COMPAT, EXACT or TUNED");
    TRSPMM_CHECK(cfg.sptc_min_nnz >= 0 && cfg.sptc_min_nnz <= 33, "sptc_min_nnz must be in [0, 33]");
    TRSPMM_CHECK((cfg.flags & ~(TRSPMM_ROUTE_EDGES | TRSPMM_ROUTE_REORDER)) == 0, "unknown route flags");
    TRSPMM_CHECK(rows >= 0 && nnz >= 0 && nnz < ((int64_t)1 << 31), "bad rows / nnz");
    TRSPMM_CHECK(ncols > 0 && ncols < ((int64_t)1 << 31), "ncols must be in [1, 2^31)");
    const int window = cfg.window;
    const int64_t W = (rows + window - 1) / window;
    const int colbits = bit_width(ncols - 1);
    const int winbits = bit_width(W > 0 ? W - 1 : 0);
    TRSPMM_CHECK(colbits + winbits <= 64, "index range too large");
    M.window = window;
    M.rows = rows;
    M.nnz = nnz;
    const bool exact = cfg.mode == TRSPMM_MODE_EXACT;
    int rc;

    // 1. validate + keys
    DevBuf err;
    if ((rc = err.alloc(A, 4))) return rc;
    TRSPMM_HIP(hipMemsetAsync(err.p, 0, 4, s));
    if (rows > 0)
        hipLaunchKernelGGL(k_validate, dim3(nblk(rows, 4)), dim3(256), 0, s, rowptr, col, rows, nnz, ncols,
                           err.as<int>());
    int herr = 0;
    TRSPMM_HIP(hipMemcpyAsync(&herr, err.p, 4, hipMemcpyDeviceToHost, s));
    TRSPMM_HIP(hipStreamSynchronize(s));
    TRSPMM_CHECK(!(herr & 1), "row pointers must start at 0, be non-decreasing and end at nnz");
    TRSPMM_CHECK(!(herr & 2), "column index out of [0, ncols)");
    TRSPMM_CHECK(!(herr & 4), "columns must be strictly ascending within each row (canonical CSR)");

    DevBuf keys, keys2, vals, svals, erow, flag, run_of, run_start, run_col, run_win, win_rb;
    const int64_t n1 = nnz > 0 ? nnz : 1;
    if ((rc = keys.alloc(A, 8 * n1)) || (rc = keys2.alloc(A, 8 * n1)) || (rc = vals.alloc(A, 4 * n1)) ||
        (rc = svals.alloc(A, 4 * n1)) || (rc = erow.alloc(A, 4 * n1)) || (rc = flag.alloc(A, 4 * n1)) ||
        (rc = run_of.alloc(A, 4 * (n1 + 1))) || (rc = win_rb.alloc(A, 4 * (W + 1))))
        return rc;
    if (nnz > 0) {
        hipLaunchKernelGGL(k_keys, dim3(nblk_gs(nnz)), dim3(256), 0, s, rowptr, col, rows, nnz, window, colbits,
                           keys.as<uint64_t>(), vals.as<int32_t>(), erow.as<int32_t>());
        TRSPMM_HIP(hipGetLastError());
        // 2. stable radix sort by (window, column)
        size_t tmp_bytes = 0;
        TRSPMM_HIP(rocprim::radix_sort_pairs(nullptr, tmp_bytes, keys.as<uint64_t>(), keys2.as<uint64_t>(),
                                             vals.as<int32_t>(), svals.as<int32_t>(), (size_t)nnz, 0,
                                             colbits + winbits, s));
        DevBuf tmp;
        if ((rc = tmp.alloc(A, tmp_bytes ? tmp_bytes : 1))) return rc;
        TRSPMM_HIP(rocprim::radix_sort_pairs(tmp.p, tmp_bytes, keys.as<uint64_t>(), keys2.as<uint64_t>(),
                                             vals.as<int32_t>(), svals.as<int32_t>(), (size_t)nnz, 0,
                                             colbits + winbits, s));
        hipLaunchKernelGGL(k_run_flags, dim3(nblk_gs(nnz)), dim3(256), 0, s, keys2.as<uint64_t>(), nnz,
                           flag.as<int32_t>());
        TRSPMM_HIP(hipGetLastError());
    }
    if ((rc = exclusive_scan<int32_t>(nnz > 0 ? flag.as<int32_t>() : nullptr, run_of.as<int32_t>(), nnz, s, &A)))
        return rc;
    int32_t nruns = 0;
    TRSPMM_HIP(hipMemcpyAsync(&nruns, run_of.as<int32_t>() + nnz, 4, hipMemcpyDeviceToHost, s));
    TRSPMM_HIP(hipStreamSynchronize(s));
    const int64_t nr1 = nruns > 0 ? nruns : 1;
    if ((rc = run_start.alloc(A, 4 * (nr1 + 1))) || (rc = run_col.alloc(A, 4 * nr1)) || (rc = run_win.alloc(A, 4 * nr1)))
        return rc;
    if (nnz > 0) {
        hipLaunchKernelGGL(k_runs, dim3(nblk_gs(nnz)), dim3(256), 0, s, keys2.as<uint64_t>(), flag.as<int32_t>(),
                           run_of.as<int32_t>(), nnz, colbits, run_start.as<int32_t>(), run_col.as<int32_t>(),
                           run_win.as<int32_t>());
        TRSPMM_HIP(hipGetLastError());
    }
    {
        const int32_t tail = (int32_t)nnz;
        TRSPMM_HIP(hipMemcpyAsync(run_start.as<int32_t>() + nruns, &tail, 4, hipMemcpyHostToDevice, s));
    }
    hipLaunchKernelGGL(k_win_runs, dim3(nblk(W + 1)), dim3(256), 0, s, run_win.as<int32_t>(), (int64_t)nruns, W,
                       win_rb.as<int32_t>());
    TRSPMM_HIP(hipGetLastError());
    keys.release();

    // 3. per-window decisions
    DevBuf order, rest, tc_chunk, in_tc, units_g, used_g, pair_a, pair_b;
    DevBuf n_tcsub, n_tcval, n_pairs, n_groups, leftover, fallback, n_rest;
    if ((rc = order.alloc(A, 4 * nr1)) || (rc = rest.alloc(A, 4 * nr1)) || (rc = tc_chunk.alloc(A, 4 * nr1)) ||
        (rc = in_tc.alloc(A, nr1)) || (rc = units_g.alloc(A, 4 * nr1)) || (rc = used_g.alloc(A, nr1)) ||
        (rc = pair_a.alloc(A, 4 * nr1)) || (rc = pair_b.alloc(A, 4 * nr1)) || (rc = n_tcsub.alloc(A, 4 * (W + 1))) ||
        (rc = n_tcval.alloc(A, 4 * (W + 1))) || (rc = n_pairs.alloc(A, 4 * (W + 1))) || (rc = n_groups.alloc(A, 4 * (W + 1))) ||
        (rc = leftover.alloc(A, 4 * (W + 1))) || (rc = fallback.alloc(A, 4 * (W + 1))) || (rc = n_rest.alloc(A, 4 * (W + 1))))
        return rc;
    RunArrays R{run_start.as<int32_t>(), run_col.as<int32_t>(), win_rb.as<int32_t>(), order.as<int32_t>(),
                rest.as<int32_t>(), tc_chunk.as<int32_t>(), in_tc.as<uint8_t>(), units_g.as<uint32_t>(),
                used_g.as<uint8_t>(), pair_a.as<int32_t>(), pair_b.as<int32_t>()};
    WinOut O{n_tcsub.as<int32_t>(), n_tcval.as<int32_t>(), n_pairs.as<int32_t>(), n_groups.as<int32_t>(),
             leftover.as<int32_t>(), fallback.as<int32_t>(), n_rest.as<int32_t>()};
    if (W > 0) {
        // size classes by units-per-window: small windows share a CU, large ones take as much
        // LDS as a workgroup may hold (160 KB on gfx950), a window beyond that sorts in global
        // memory.  A unit takes 4 bytes (key) + 1 (used flag).
        // classes 1-3 run on whole workgroups (k_window_decide_wg: 4 waves for the
        // medium class, 16 for the large ones)
        int big_cap = 12288;
        {
            int dev = 0, max_lds = 0;
            hipFuncAttributes fa{};
            if (hipGetDevice(&dev) == hipSuccess &&
                hipDeviceGetAttribute(&max_lds, hipDeviceAttributeMaxSharedMemoryPerBlock, dev) == hipSuccess &&
                hipFuncGetAttributes(&fa, (const void*)k_window_decide_wg<16>) == hipSuccess) {
                const int cap = (int)((max_lds - (int64_t)fa.sharedSizeBytes - 64) / 5) & ~1;
                if (cap > big_cap &&
                    hipFuncSetAttribute((const void*)k_window_decide_wg<16>,
                                        hipFuncAttributeMaxDynamicSharedMemorySize, cap * 5 + 16) == hipSuccess)
                    big_cap = cap;
            }
            (void)hipGetLastError();      // a refused query leaves the 64 KB classes
        }
        // windows by size class (units <= ceil(columns / 2); TC columns only reduce it), each
        // class listed largest first so the longest windows start first
        std::vector<int32_t> rb_h((size_t)W + 1);
        TRSPMM_HIP(hipMemcpyAsync(rb_h.data(), win_rb.as<int32_t>(), 4 * (W + 1), hipMemcpyDeviceToHost, s));
        TRSPMM_HIP(hipStreamSynchronize(s));
        const int caps[4] = {1024, std::min(6144, big_cap), big_cap, 0};
        std::vector<int32_t> lists[4];
        for (int64_t w = 0; w < W; ++w) {
            const int32_t umax = (rb_h[w + 1] - rb_h[w] + 1) / 2;
            const int c = umax <= caps[0] ? 0 : umax <= caps[1] ? 1 : umax <= caps[2] ? 2 : 3;
            lists[c].push_back((int32_t)w);
        }
        std::vector<int32_t> all;
        size_t at[5] = {0, 0, 0, 0, 0};
        for (int c = 0; c < 4; ++c) {
            if (c > 0)
                std::stable_sort(lists[c].begin(), lists[c].end(), [&](int32_t a, int32_t b) {
                    return rb_h[a + 1] - rb_h[a] > rb_h[b + 1] - rb_h[b];
                });
            all.insert(all.end(), lists[c].begin(), lists[c].end());
            at[c + 1] = all.size();
        }
        DevBuf wl;
        if ((rc = wl.alloc(A, 4 * all.size()))) return rc;
        TRSPMM_HIP(hipMemcpyAsync(wl.p, all.data(), 4 * all.size(), hipMemcpyHostToDevice, s));
        // the classes are independent: the three smaller ones run on side streams
        // while the largest (windows sorted in global memory) runs on s.  (On
        // R-MAT scale 18 with reordered rows classes 2 and 3 each hold a hub
        // window of several ms: one after the other on s they took 10.6 ms.)
        hipStream_t side[kSide] = {s, s, s};
        EventSet ev;
        const bool fork = (at[1] > at[0]) + (at[2] > at[1]) + (at[3] > at[2]) + (at[4] > at[3]) > 1;
        if (fork) {
            if ((rc = side_streams(side)) || (rc = ev.make())) return rc;
            TRSPMM_HIP(hipEventRecord(ev.e[0], s));
            for (int i = 0; i < kSide; ++i) {
                ev.side[i] = side[i];
                TRSPMM_HIP(hipStreamWaitEvent(side[i], ev.e[0], 0));
            }
        }
        for (int c = 3; c >= 0; --c) {
            const size_t n = at[c + 1] - at[c];
            if (n == 0) continue;
            const size_t lds = caps[c] > 0 ? (size_t)caps[c] * 5 + 16 : 0;
            if (c == 0)
                hipLaunchKernelGGL(k_window_decide, dim3((unsigned)n), dim3(64), lds, side[0], R, O, W, window,
                                   cfg.density, cfg.tc_min_util, cfg.skip_tc, cfg.unit_min_nnz,
                                   wl.as<int32_t>() + at[c], caps[c]);
            else if (c == 1)
                hipLaunchKernelGGL(k_window_decide_wg<4>, dim3((unsigned)n), dim3(4 * kWave), lds, side[1], R, O, W,
                                   window, cfg.density, cfg.tc_min_util, cfg.skip_tc, cfg.unit_min_nnz,
                                   wl.as<int32_t>() + at[c], caps[c], 0);
            else
                hipLaunchKernelGGL(k_window_decide_wg<16>, dim3((unsigned)n), dim3(16 * kWave), lds,
                                   c == 2 ? side[2] : s, R, O, W, window, cfg.density, cfg.tc_min_util, cfg.skip_tc,
                                   cfg.unit_min_nnz, wl.as<int32_t>() + at[c], caps[c], kHostPairUnits);
            TRSPMM_HIP(hipGetLastError());
        }
        if (fork) {
            for (int i = 0; i < kSide; ++i) {
                TRSPMM_HIP(hipEventRecord(ev.e[1 + i], side[i]));
                TRSPMM_HIP(hipStreamWaitEvent(s, ev.e[1 + i], 0));
            }
            ev.joined = true;
        }
        TRSPMM_HIP(hipStreamSynchronize(s));
        // windows of classes 2 and 3 with more than kHostPairUnits units: the
        // pairing on the host, their keys packed into one download and their
        // pairs into one upload
        std::vector<int32_t> hostw;
        for (size_t k = at[2]; k < at[4]; ++k)
            if ((rb_h[all[k] + 1] - rb_h[all[k]] + 1) / 2 > kHostPairUnits) hostw.push_back(all[k]);
        std::vector<int32_t> nrest_h;
        if (!hostw.empty()) {
            nrest_h.resize((size_t)W);
            TRSPMM_HIP(hipMemcpyAsync(nrest_h.data(), O.n_rest, 4 * W, hipMemcpyDeviceToHost, s));
            TRSPMM_HIP(hipStreamSynchronize(s));
            size_t keep = 0;                           // the windows whose units exceed the threshold
            for (int32_t w : hostw)
                if ((nrest_h[w] + 1) / 2 > kHostPairUnits) hostw[keep++] = w;
            hostw.resize(keep);
        }
        const int64_t nbig = (int64_t)hostw.size();
        if (nbig > 0) {
            // per window: [0] the window, [1] its first packed unit, [2] pairs, [3] leftover
            std::vector<int32_t> tab((size_t)(4 * nbig));
            int64_t total = 0;
            for (int64_t b = 0; b < nbig; ++b) {
                tab[4 * b] = hostw[b];
                tab[4 * b + 1] = (int32_t)total;
                total += (nrest_h[hostw[b]] + 1) / 2;
            }
            TRSPMM_CHECK(total < ((int64_t)1 << 31), "router: too many units in the hub windows");
            DevBuf tab_d, keys_d, pa_d, pb_d;
            if ((rc = tab_d.alloc(A, 16 * nbig)) || (rc = keys_d.alloc(A, 4 * total)) || (rc = pa_d.alloc(A, 4 * total)) ||
                (rc = pb_d.alloc(A, 4 * total)))
                return rc;
            TRSPMM_HIP(hipMemcpyAsync(tab_d.p, tab.data(), 16 * nbig, hipMemcpyHostToDevice, s));
            hipLaunchKernelGGL(k_pack_units, dim3((unsigned)nbig), dim3(256), 0, s, R, O, tab_d.as<int32_t>(),
                               keys_d.as<uint32_t>());
            TRSPMM_HIP(hipGetLastError());
            std::vector<uint32_t> keys((size_t)total);
            std::vector<int32_t> pa((size_t)total), pb((size_t)total);
            TRSPMM_HIP(hipMemcpyAsync(keys.data(), keys_d.p, 4 * total, hipMemcpyDeviceToHost, s));
            TRSPMM_HIP(hipStreamSynchronize(s));
            for (int64_t b = 0; b < nbig; ++b) {
                const int32_t off = tab[4 * b + 1];
                const int32_t nu = (nrest_h[hostw[b]] + 1) / 2;
                int32_t np = 0, left_k = nu;
                if (nu > 0) pair_on_host(keys.data() + off, nu, pa.data() + off, pb.data() + off, np, left_k);
                tab[4 * b + 2] = np;
                tab[4 * b + 3] = left_k;
            }
            TRSPMM_HIP(hipMemcpyAsync(tab_d.p, tab.data(), 16 * nbig, hipMemcpyHostToDevice, s));
            TRSPMM_HIP(hipMemcpyAsync(pa_d.p, pa.data(), 4 * total, hipMemcpyHostToDevice, s));
            TRSPMM_HIP(hipMemcpyAsync(pb_d.p, pb.data(), 4 * total, hipMemcpyHostToDevice, s));
            hipLaunchKernelGGL(k_window_finish, dim3((unsigned)nbig), dim3(256), 0, s, R, O, tab_d.as<int32_t>(),
                               pa_d.as<int32_t>(), pb_d.as<int32_t>(), cfg.unit_min_nnz);
            TRSPMM_HIP(hipGetLastError());
            TRSPMM_HIP(hipStreamSynchronize(s));    // the host vectors are in the copies above
        }
    }

    // 4. window prefix sums -> output sizes
    DevBuf n_tparts, n_tiles, tc_sub_off, tc_val_off, tc_part_off, grp_all_off;
    if ((rc = n_tparts.alloc(A, 4 * (W + 1))) || (rc = n_tiles.alloc(A, 4 * (W + 1))) ||
        (rc = tc_sub_off.alloc(A, 4 * (W + 1))) || (rc = tc_val_off.alloc(A, 4 * (W + 1))) ||
        (rc = tc_part_off.alloc(A, 4 * (W + 1))) || (rc = grp_all_off.alloc(A, 4 * (W + 1))))
        return rc;
    if ((rc = M.alloc(22, W + 1)) || (rc = M.alloc(28, W + 1))) return rc;   // s_offsets, s_window_tile_offset
    if (W > 0) {
        hipLaunchKernelGGL(k_part_counts, dim3(nblk(W)), dim3(256), 0, s, n_tcsub.as<int32_t>(),
                           n_groups.as<int32_t>(), W, cfg.part_t, n_tparts.as<int32_t>(), n_tiles.as<int32_t>());
        TRSPMM_HIP(hipGetLastError());
    }
    const int32_t* nz = nullptr;
    if ((rc = exclusive_scan<int32_t>(W ? n_tcsub.as<int32_t>() : nz, tc_sub_off.as<int32_t>(), W, s, &A)) ||
        (rc = exclusive_scan<int32_t>(W ? n_tcval.as<int32_t>() : nz, tc_val_off.as<int32_t>(), W, s, &A)) ||
        (rc = exclusive_scan<int32_t>(W ? n_tparts.as<int32_t>() : nz, tc_part_off.as<int32_t>(), W, s, &A)) ||
        (rc = exclusive_scan<int32_t>(W ? n_groups.as<int32_t>() : nz, grp_all_off.as<int32_t>(), W, s, &A)))
        return rc;
    int32_t tot[5];
    TRSPMM_HIP(hipMemcpyAsync(&tot[0], tc_sub_off.as<int32_t>() + W, 4, hipMemcpyDeviceToHost, s));
    TRSPMM_HIP(hipMemcpyAsync(&tot[1], tc_val_off.as<int32_t>() + W, 4, hipMemcpyDeviceToHost, s));
    TRSPMM_HIP(hipMemcpyAsync(&tot[2], tc_part_off.as<int32_t>() + W, 4, hipMemcpyDeviceToHost, s));
    TRSPMM_HIP(hipMemcpyAsync(&tot[3], grp_all_off.as<int32_t>() + W, 4, hipMemcpyDeviceToHost, s));
    TRSPMM_HIP(hipStreamSynchronize(s));
    const int64_t S = tot[0], TV = tot[1], PT = tot[2], GA = tot[3];

    // 2:4 groups that stay on the 2:4 route (all but sparse ones in TUNED mode)
    DevBuf res_flag, keep, gmap;
    if ((rc = res_flag.alloc(A, n1)) || (rc = keep.alloc(A, GA > 0 ? GA : 1)) || (rc = gmap.alloc(A, 4 * (GA + 1))))
        return rc;
    TRSPMM_HIP(hipMemsetAsync(res_flag.p, 0, n1, s));
    Emit E{svals.as<int32_t>(), erow.as<int32_t>(), val, tc_sub_off.as<int32_t>(), tc_val_off.as<int32_t>(),
           tc_part_off.as<int32_t>(), (const int32_t*)M.data[22], (const int32_t*)M.data[28],
           grp_all_off.as<int32_t>(), gmap.as<int32_t>(), keep.as<uint8_t>()};
    if (GA > 0) {
        hipLaunchKernelGGL(k_group_keep, dim3(nblk(GA)), dim3(256), 0, s, R, O, E, W, window,
                           cfg.mode == TRSPMM_MODE_TUNED ? cfg.sptc_min_nnz : 0, GA);
        TRSPMM_HIP(hipGetLastError());
    }
    // gmap = exclusive scan of keep (bytes widened through the int32 counts buffer)
    {
        DevBuf keep32;
        if ((rc = keep32.alloc(A, 4 * (GA > 0 ? GA : 1)))) return rc;
        if (GA > 0) {
            hipLaunchKernelGGL(k_widen, dim3(nblk(GA)), dim3(256), 0, s, keep.as<uint8_t>(), GA, keep32.as<int32_t>());
            TRSPMM_HIP(hipGetLastError());
        }
        if ((rc = exclusive_scan<int32_t>(GA ? keep32.as<int32_t>() : nz, gmap.as<int32_t>(), GA, s, &A))) return rc;
    }
    hipLaunchKernelGGL(k_group_offsets, dim3(nblk(W + 1)), dim3(256), 0, s, grp_all_off.as<int32_t>(),
                       gmap.as<int32_t>(), W, (int32_t*)M.data[22], n_tiles.as<int32_t>());
    TRSPMM_HIP(hipGetLastError());
    if ((rc = exclusive_scan<int32_t>(W ? n_tiles.as<int32_t>() : nz, (int32_t*)M.data[28], W, s, &A))) return rc;
    TRSPMM_HIP(hipMemcpyAsync(&tot[3], (int32_t*)M.data[22] + W, 4, hipMemcpyDeviceToHost, s));
    TRSPMM_HIP(hipMemcpyAsync(&tot[4], (int32_t*)M.data[28] + W, 4, hipMemcpyDeviceToHost, s));
    TRSPMM_HIP(hipStreamSynchronize(s));
    const int64_t G = tot[3], T = tot[4];

    // 5. outputs
    if ((rc = M.alloc(0, PT + 1)) || (rc = M.alloc(1, S + 1)) || (rc = M.alloc(2, 8 * S)) || (rc = M.alloc(3, TV)) ||
        (rc = M.alloc(4, PT)) || (rc = M.alloc(5, PT)) || (rc = M.alloc(6, 2 * S)) || (rc = M.alloc(17, 4 * G)) ||
        (rc = M.alloc(18, G * window)) || (rc = M.alloc(19, G * window * 2)) || (rc = M.alloc(20, G * window * 2)) ||
        (rc = M.alloc(21, G)) || (rc = M.alloc(23, T)) || (rc = M.alloc(24, T)) || (rc = M.alloc(25, 16 * T)) ||
        (rc = M.alloc(26, 64 * T)) || (rc = M.alloc(27, 8 * T)))
        return rc;
    E.t_row_offset = (int32_t*)M.data[0];
    E.t_block_offset = (int32_t*)M.data[1];
    E.t_column = (int32_t*)M.data[2];
    E.t_value = (float*)M.data[3];
    E.t_window_row = (int32_t*)M.data[4];
    E.t_atomic = (int32_t*)M.data[5];
    E.t_binary = (int64_t*)M.data[6];
    E.s_column = (int32_t*)M.data[17];
    E.s_mask = (uint8_t*)M.data[18];
    E.s_value = (float*)M.data[19];
    E.s_pos = (int8_t*)M.data[20];
    E.s_window = (int32_t*)M.data[21];
    E.s_offsets = (int32_t*)M.data[22];
    E.s_tile_window = (int32_t*)M.data[23];
    E.s_tile_group = (int32_t*)M.data[24];
    E.s_tile_column = (int32_t*)M.data[25];
    E.s_packed_a = (int32_t*)M.data[26];
    E.s_packed_meta = (int32_t*)M.data[27];
    E.s_window_tile_offset = (int32_t*)M.data[28];
    E.res_flag = res_flag.as<uint8_t>();
    if (cfg.flags & TRSPMM_ROUTE_EDGES) {
        if ((rc = M.edge_alloc(0, TV)) || (rc = M.edge_alloc(3, G * window * 2))) return rc;
        E.t_edge = (int32_t*)M.edge_data[0];
        E.s_edge = (int32_t*)M.edge_data[3];
        if (TV > 0) TRSPMM_HIP(hipMemsetAsync(E.t_edge, 0xff, 4 * TV, s));   // -1: no edge
    }
    if (S > 0) {
        DevBuf tcn;
        if ((rc = tcn.alloc(A, 4 * (size_t)S))) return rc;
        hipLaunchKernelGGL(k_tc_counts, dim3(nblk(S)), dim3(256), 0, s, R, E, W, S, tcn.as<int32_t>());
        TRSPMM_HIP(hipGetLastError());
        if ((rc = exclusive_scan<int32_t>(tcn.as<int32_t>(), E.t_block_offset, S, s, &A))) return rc;
        hipLaunchKernelGGL(k_emit_tc, dim3(nblk(S)), dim3(256), 0, s, R, O, E, W, window, cfg.part_t, S);
        TRSPMM_HIP(hipGetLastError());
    }
    hipLaunchKernelGGL(k_tail, dim3(1), dim3(1), 0, s, E, S, PT, TV);
    if (GA > 0) {
        hipLaunchKernelGGL(k_emit_groups, dim3(nblk(GA)), dim3(256), 0, s, R, O, E, W, window,
                           exact || cfg.mode == TRSPMM_MODE_TUNED ? 1 : 0, GA);
        TRSPMM_HIP(hipGetLastError());
    }
    if (W > 0) {
        hipLaunchKernelGGL(k_flag_dropped, dim3(nblk(W)), dim3(256), 0, s, R, O, E, W);
        TRSPMM_HIP(hipGetLastError());
    }
    if (T > 0) {
        hipLaunchKernelGGL(k_pack_tiles, dim3(nblk(T * 8)), dim3(256), 0, s, E, W, window, T);
        TRSPMM_HIP(hipGetLastError());
    }

    // residual parts per row
    DevBuf nl, ns, el, es, pl, ps, ql, qs;
    const int64_t rw1 = rows + 1;
    if ((rc = nl.alloc(A, 4 * rw1)) || (rc = ns.alloc(A, 4 * rw1)) || (rc = el.alloc(A, 4 * rw1)) || (rc = es.alloc(A, 4 * rw1)) ||
        (rc = pl.alloc(A, 4 * rw1)) || (rc = ps.alloc(A, 4 * rw1)) || (rc = ql.alloc(A, 4 * rw1)) || (rc = qs.alloc(A, 4 * rw1)))
        return rc;
    ResCounts RCn{nl.as<int32_t>(), ns.as<int32_t>(), el.as<int32_t>(), es.as<int32_t>()};
    if (rows > 0) {
        hipLaunchKernelGGL(k_res_count, dim3(nblk(rows, 4)), dim3(256), 0, s, rowptr, res_flag.as<uint8_t>(), rows,
                           cfg.part_c, cfg.short_c, RCn);
        TRSPMM_HIP(hipGetLastError());
    }
    const int32_t* z = nullptr;
    if ((rc = exclusive_scan<int32_t>(rows ? nl.as<int32_t>() : z, pl.as<int32_t>(), rows, s, &A)) ||
        (rc = exclusive_scan<int32_t>(rows ? ns.as<int32_t>() : z, ps.as<int32_t>(), rows, s, &A)) ||
        (rc = exclusive_scan<int32_t>(rows ? el.as<int32_t>() : z, ql.as<int32_t>(), rows, s, &A)) ||
        (rc = exclusive_scan<int32_t>(rows ? es.as<int32_t>() : z, qs.as<int32_t>(), rows, s, &A)))
        return rc;
    int32_t rt[4];
    TRSPMM_HIP(hipMemcpyAsync(&rt[0], pl.as<int32_t>() + rows, 4, hipMemcpyDeviceToHost, s));
    TRSPMM_HIP(hipMemcpyAsync(&rt[1], ps.as<int32_t>() + rows, 4, hipMemcpyDeviceToHost, s));
    TRSPMM_HIP(hipMemcpyAsync(&rt[2], ql.as<int32_t>() + rows, 4, hipMemcpyDeviceToHost, s));
    TRSPMM_HIP(hipMemcpyAsync(&rt[3], qs.as<int32_t>() + rows, 4, hipMemcpyDeviceToHost, s));
    TRSPMM_HIP(hipStreamSynchronize(s));
    const int64_t PL = rt[0], PS = rt[1], EL = rt[2], ES = rt[3];
    if ((rc = M.alloc(7, PL + 1)) || (rc = M.alloc(8, PL)) || (rc = M.alloc(9, PL)) || (rc = M.alloc(10, EL)) ||
        (rc = M.alloc(11, EL)) || (rc = M.alloc(12, PS + 1)) || (rc = M.alloc(13, PS)) || (rc = M.alloc(14, PS)) ||
        (rc = M.alloc(15, ES)) || (rc = M.alloc(16, ES)))
        return rc;
    ResOut RO{pl.as<int32_t>(), ps.as<int32_t>(), ql.as<int32_t>(), qs.as<int32_t>(),
              (int32_t*)M.data[7], (int32_t*)M.data[8], (int32_t*)M.data[9], (int32_t*)M.data[10],
              (float*)M.data[11], (int32_t*)M.data[12], (int32_t*)M.data[13], (int32_t*)M.data[14],
              (int32_t*)M.data[15], (float*)M.data[16]};
    if (cfg.flags & TRSPMM_ROUTE_EDGES) {
        if ((rc = M.edge_alloc(1, EL)) || (rc = M.edge_alloc(2, ES))) return rc;
        RO.c_edge = (int32_t*)M.edge_data[1];
        RO.cs_edge = (int32_t*)M.edge_data[2];
    }
    if (rows > 0) {
        hipLaunchKernelGGL(k_res_emit, dim3(nblk(rows, 4)), dim3(256), 0, s, rowptr, col, val, res_flag.as<uint8_t>(),
                           n_tcsub.as<int32_t>(), rows, window, cfg.part_c, cfg.short_c, RO);
        TRSPMM_HIP(hipGetLastError());
    }
    hipLaunchKernelGGL(k_res_tail, dim3(1), dim3(1), 0, s, RO, rows, PL, PS);
    TRSPMM_HIP(hipGetLastError());
    TRSPMM_HIP(hipStreamSynchronize(s));
    const auto t_end = std::chrono::steady_clock::now();
    M.seconds = std::chrono::duration<double>(t_end - t_begin).count();
    return TRSPMM_OK;
}

namespace {

// B rows one multiply over a routing gathers: TC sub-block columns, 2:4 tile
// columns and residual entries (the quantity the reordering is judged by)
int gathered_rows(const Meta& M, hipStream_t s, StreamArena& A, int64_t* out) {
    DevBuf c;
    int rc;
    if ((rc = c.alloc(A, 8))) return rc;
    TRSPMM_HIP(hipMemsetAsync(c.p, 0, 8, s));
    const int idx[2] = {2, 25};          // t_column [8 S], s_tile_column [16 T]
    for (int i : idx)
        if (M.count[i] > 0) {
            hipLaunchKernelGGL(k_count_valid, dim3(nblk_gs(M.count[i])), dim3(256), 0, s, (const int32_t*)M.data[i],
                               M.count[i], c.as<unsigned long long>());
            TRSPMM_HIP(hipGetLastError());
        }
    unsigned long long h = 0;
    TRSPMM_HIP(hipMemcpyAsync(&h, c.p, 8, hipMemcpyDeviceToHost, s));
    TRSPMM_HIP(hipStreamSynchronize(s));
    *out = (int64_t)h + M.count[10] + M.count[15];     // + c_column, c_column_short
    return TRSPMM_OK;
}

}  // namespace

// a second stream per device, on which the input-order routing runs beside a
// reordered one
static int aux_stream(hipStream_t* out) {
    static std::mutex mu;
    static std::map<int, hipStream_t> made;
    int dev = 0;
    TRSPMM_HIP(hipGetDevice(&dev));
    std::lock_guard<std::mutex> g(mu);
    auto it = made.find(dev);
    if (it == made.end()) {
        hipStream_t a = nullptr;
        TRSPMM_HIP(hipStreamCreateWithFlags(&a, hipStreamNonBlocking));
        it = made.emplace(dev, a).first;
    }
    *out = it->second;
    return TRSPMM_OK;
}

// the input-order routing of a reordered route, on a host thread of its own
// (route_core waits for the device between its stages) and stream s2 after the
// event `ready` of the caller's stream; its smaller size classes share the side
// streams with the reordered routing (each fork waits for its own events only).
// Joined by the destructor on every path
struct NaturalRoute {
    std::thread t;
    int rc = TRSPMM_OK;
    std::string err;
    void start(const int32_t* rowptr, const int32_t* col, const float* val, int64_t rows, int64_t nnz,
               int64_t ncols, const trspmm_route_cfg& cfg, hipStream_t s2, hipEvent_t ready, int dev, Meta& M) {
        t = std::thread([=, &M] {
            if (hipSetDevice(dev) != hipSuccess || hipStreamWaitEvent(s2, ready, 0) != hipSuccess) {
                (void)hipGetLastError();
                rc = TRSPMM_EHIP;
                err = "input-order routing: hipSetDevice / hipStreamWaitEvent failed";
                return;
            }
            try {          // an exception must not leave the thread (std::terminate)
                rc = route_core(rowptr, col, val, rows, nnz, ncols, cfg, s2, M);
            } catch (const std::exception& e) {
                rc = TRSPMM_ENOMEM;
                set_error(std::string("input-order routing: ") + e.what());
            }
            if (hipStreamSynchronize(s2) != hipSuccess && rc == TRSPMM_OK) {
                (void)hipGetLastError();
                rc = TRSPMM_EHIP;
                set_error("input-order routing: hipStreamSynchronize failed");
            }
            if (rc != TRSPMM_OK) err = last_error();
        });
    }
    void join() {
        if (t.joinable()) t.join();
    }
    ~NaturalRoute() { join(); }
};

int route_device(const int32_t* rowptr, const int32_t* col, const float* val, int64_t rows, int64_t nnz,
                 int64_t ncols, const trspmm_route_cfg& cfg, hipStream_t s, Meta& M) {
    const auto t_begin = std::chrono::steady_clock::now();
    const bool reorder = (cfg.flags & TRSPMM_ROUTE_REORDER) != 0;
    TRSPMM_CHECK(!reorder || (cfg.mode != TRSPMM_MODE_COMPAT && cfg.window == 16),
                 "row reordering needs EXACT or TUNED mode and window 16");
    if (!reorder || rows < 2 || nnz == 0) return route_core(rowptr, col, val, rows, nnz, ncols, cfg, s, M);

    // Both row orders are routed, at the same time: the input order on a second
    // stream from a second host thread, the reordered rows here.  The reordered
    // routing is kept when it gathers at least a tenth fewer B rows.  The count
    // does not see what the input order's windows share through the caches
    // beyond one window (communities and bands spanning windows): on the
    // three-route mix (2^20 rows, K = 256) 1.9 % fewer rows ran 11 % slower, on
    // R-MAT scale 18 14 % fewer rows ran 10 % faster.
    int rc;
    hipStream_t s2 = nullptr;
    int dev = 0;
    if ((rc = aux_stream(&s2))) return rc;
    TRSPMM_HIP(hipGetDevice(&dev));
    struct Event {
        hipEvent_t e = nullptr;
        ~Event() {
            if (e) (void)hipEventDestroy(e);
        }
    } ready;
    TRSPMM_HIP(hipEventCreateWithFlags(&ready.e, hipEventDisableTiming));
    TRSPMM_HIP(hipEventRecord(ready.e, s));
    StreamArena A(s);
    Meta R;
    NaturalRoute nat;                          // declared after what its thread uses: joined first
    nat.start(rowptr, col, val, rows, nnz, ncols, cfg, s2, ready.e, dev, M);
    DevBuf keys, keys2, idx, order, cnt, prowptr, pcol, pval, pedge, tmp;
    if ((rc = keys.alloc(A, 8 * rows)) || (rc = keys2.alloc(A, 8 * rows)) || (rc = idx.alloc(A, 4 * rows)) ||
        (rc = order.alloc_cached(4 * rows)) || (rc = cnt.alloc(A, 4 * rows)) || (rc = prowptr.alloc(A, 4 * (rows + 1))) ||
        (rc = pcol.alloc(A, 4 * nnz)) || (rc = pval.alloc(A, 4 * nnz)) || (rc = pedge.alloc(A, 4 * nnz)))
        return rc;
    hipLaunchKernelGGL(k_row_keys, dim3(nblk_gs(rows)), dim3(256), 0, s, rowptr, col, rows, keys.as<uint64_t>(),
                       idx.as<int32_t>());
    TRSPMM_HIP(hipGetLastError());
    size_t tmp_bytes = 0;
    TRSPMM_HIP(rocprim::radix_sort_pairs(nullptr, tmp_bytes, keys.as<uint64_t>(), keys2.as<uint64_t>(),
                                         idx.as<int32_t>(), order.as<int32_t>(), (size_t)rows, 0, 64, s));
    if ((rc = tmp.alloc(A, tmp_bytes ? tmp_bytes : 1))) return rc;
    TRSPMM_HIP(rocprim::radix_sort_pairs(tmp.p, tmp_bytes, keys.as<uint64_t>(), keys2.as<uint64_t>(),
                                         idx.as<int32_t>(), order.as<int32_t>(), (size_t)rows, 0, 64, s));
    hipLaunchKernelGGL(k_perm_counts, dim3(nblk_gs(rows)), dim3(256), 0, s, rowptr, order.as<int32_t>(), rows,
                       cnt.as<int32_t>());
    TRSPMM_HIP(hipGetLastError());
    if ((rc = exclusive_scan<int32_t>(cnt.as<int32_t>(), prowptr.as<int32_t>(), rows, s, &A))) return rc;
    hipLaunchKernelGGL(k_perm_rows, dim3(nblk(rows, 4)), dim3(256), 0, s, rowptr, col, val, order.as<int32_t>(),
                       prowptr.as<int32_t>(), rows, pcol.as<int32_t>(), pval.as<float>(), pedge.as<int32_t>());
    TRSPMM_HIP(hipGetLastError());
    trspmm_route_cfg c2 = cfg;
    c2.flags &= ~TRSPMM_ROUTE_REORDER;
    if ((rc = route_core(prowptr.as<int32_t>(), pcol.as<int32_t>(), pval.as<float>(), rows, nnz, ncols, c2, s, R)))
        return rc;
    nat.join();
    if (nat.rc != TRSPMM_OK) {
        set_error(nat.err);
        return nat.rc;
    }
    int64_t g0 = 0, g1 = 0;
    if ((rc = gathered_rows(M, s, A, &g0)) || (rc = gathered_rows(R, s, A, &g1))) return rc;
    if (10 * g1 <= 9 * g0) {
        for (int i = 0; i < 4; ++i)
            if (R.edge_count[i] > 0) {
                hipLaunchKernelGGL(k_map_edges, dim3(nblk_gs(R.edge_count[i])), dim3(256), 0, s,
                                   (int32_t*)R.edge_data[i], R.edge_count[i], pedge.as<int32_t>());
                TRSPMM_HIP(hipGetLastError());
            }
        R.row_order_buf.swap(order);
        R.row_order = R.row_order_buf.as<int32_t>();
        M.swap(R);
    }
    TRSPMM_HIP(hipStreamSynchronize(s));
    M.seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - t_begin).count();
    return TRSPMM_OK;
}

trspmm_meta_view meta_view(const Meta& M) {
    trspmm_meta_view v{};
    v.t_row_offset = (const int32_t*)M.data[0];
    v.n_tparts = M.count[4];
    v.t_block_offset = (const int32_t*)M.data[1];
    v.t_column = (const int32_t*)M.data[2];
    v.n_tsub = M.count[2] / 8;
    v.t_value = M.data[3];
    v.n_tval = M.count[3];
    v.t_window_row = (const int32_t*)M.data[4];
    v.t_binary = (const int64_t*)M.data[6];
    v.c_row_offset = (const int32_t*)M.data[7];
    v.n_cparts = M.count[8];
    v.c_row = (const int32_t*)M.data[8];
    v.c_column = (const int32_t*)M.data[10];
    v.n_cval = M.count[10];
    v.c_value = M.data[11];
    v.cs_row_offset = (const int32_t*)M.data[12];
    v.n_csparts = M.count[13];
    v.cs_row = (const int32_t*)M.data[13];
    v.cs_column = (const int32_t*)M.data[15];
    v.n_csval = M.count[15];
    v.cs_value = M.data[16];
    v.s_column = (const int32_t*)M.data[17];
    v.n_groups = M.count[21];
    v.s_value = (const float*)M.data[19];
    v.s_pos = (const int8_t*)M.data[20];
    v.s_offsets = (const int32_t*)M.data[22];
    v.n_windows = M.count[22] - 1;
    v.t_value_is_f32 = 1;
    v.c_value_is_f32 = 1;
    v.window = M.window;
    return v;
}

EdgeMaps meta_edges(const Meta& M) {
    return EdgeMaps{(const int32_t*)M.edge_data[0], (const int32_t*)M.edge_data[1], (const int32_t*)M.edge_data[2],
                    (const int32_t*)M.edge_data[3], M.nnz};
}

}  // namespace trspmm
