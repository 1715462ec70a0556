// Device statistics of the reference's matching_utils_online.match_2to4_online
// (TR-source/Block/block_matching_online.cu:85-649): the sliding-window 1:2 /
// 2:4 matching with dense-pool routing, computed on the GPU, one 64-lane
// workgroup per 16-row window, the windows launched heaviest first.
//
// Per window (rows r0 .. r0+15):
//   1. table: the reference puts the window's distinct columns into a
//      2048-slot table at slot col mod 2048 with linear probing, each with
//      the 16-bit mask of the window rows it touches, and takes them in slot
//      order.  It inserts from all threads at once, so when two columns of a
//      window share a slot the order depends on thread scheduling.  Here the
//      order is fixed as the one sequential insertion in ascending column
//      order gives (one of the reference's possible outcomes); without such a
//      collision every column sits at its own slot and the order is the
//      reference's, whatever the scheduling.  The columns are collected in a
//      scattered-hash table of 4096 slots (at most half full, so probes stay
//      short whatever the column ids), then laid out: each at its home slot
//      when no two share one, else by inserting them in ascending order, the
//      wave finding each one's first free slot in a register bitmap of the
//      table with one ballot;
//   2. 1:2 pairs: each unpaired column i takes, among columns i+1 .. i+32 not
//      yet paired, the first with no common row, else the first with the
//      fewest common rows if <= t_max; else it goes to the pool (fine);
//   3. 2:4 groups: each unused pair takes the first unused pair among the next
//      1024 whose rows do not conflict ((both_i & any_j) | (any_i & both_j)
//      == 0); a group adds 32 - nnz padding zeros; a pair with no partner
//      sends both columns to the pool (coarse 2);
//   4. the pool, in order: blocks of dense_threshold columns (TC), the rest
//      two at a time in padded 2:4 groups (32 - nnz zeros);
//   5. 2:4 groups packed four to a 16x16 block (ceil), shortfall = padding.
// A window with more than 2048 distinct columns overflows the table: the
// reference then drops the columns that find no slot, which ones depending on
// scheduling; here the window keeps its 2048 smallest column ids (what
// sequential insertion in ascending order keeps) and is counted in the
// overflow counter.
#include <hip/hip_runtime.h>

#include <cstdint>
#include <mutex>
#include <string>

#include "common.h"

namespace trspmm {

namespace {

constexpr int OM_ROWS = 16;           // window height (fixed, as in the reference kernel)
constexpr int OM_SLOTS = 2048;        // the reference's table slots per window
constexpr int OM_DEDUP = 4096;        // collection table slots (at most half full)
constexpr int OM_COUNT = 8192;        // counting table slots (ids only; at most half full)
constexpr int OM_BATCH = 4;           // column loads in flight per lane in a collection pass
constexpr int OM_PAIR_SEARCH = 32;    // columns looked ahead for a 1:2 partner
constexpr int OM_GROUP_SEARCH = 1024; // pairs looked ahead for a 2:4 partner
constexpr int OM_WAVE = 64;
constexpr uint16_t OM_FREE = 0xFFFF;

struct OnlineLds {
    union {
        struct {                          // collection: column id (-1 free) and row mask
            int32_t id[OM_DEDUP];
            uint32_t mask[OM_DEDUP];
        } d;
        int32_t cnt[OM_COUNT];            // counting passes: ids only, twice the slots
        struct {                          // layout in the reference's table, then its order
            uint32_t home_cnt[OM_SLOTS];  // columns whose home is this slot
            uint16_t owner[OM_SLOTS];     // column (index into keys) in this slot, OM_FREE if none
            uint16_t cmask[OM_SLOTS];     // row masks in table order
        } t;
    } a;
    union {
        uint32_t ids[2 * OM_SLOTS];       // cut-off search: the distinct ids below a cut, sorted
        uint64_t keys[OM_SLOTS];          // collected columns: (column id << 32) | row mask
        struct {
            // pair: rows both columns touch | rows either touches << 16; all
            // ones once the pair is in a 2:4 group
            uint32_t pk[OM_SLOTS / 2];
            uint16_t c1[OM_SLOTS / 2], c2[OM_SLOTS / 2];   // pair: column positions in table order
            uint16_t pool[OM_SLOTS];
        } p;
    } b;
};

struct OnlineCtl {
    int32_t rp[OM_ROWS + 1];      // the window's row offsets, clamped to [0, nnz]
    int32_t occ;                  // distinct columns collected
    int32_t stop;                 // the collection reached its cap
    int32_t collide;              // two collected columns share a home slot
};

__device__ inline int lane_id() { return (int)threadIdx.x; }

// exclusive prefix sum over the wave; *total = sum of all lanes
__device__ inline int wave_exclusive_scan(int v, int* total) {
    int x = v;
#pragma unroll
    for (int d = 1; d < OM_WAVE; d <<= 1) {
        const int y = __shfl_up(x, d, OM_WAVE);
        if (lane_id() >= d) x += y;
    }
    *total = __shfl(x, OM_WAVE - 1, OM_WAVE);
    return x - v;
}

__device__ inline int wave_sum(int v) {
#pragma unroll
    for (int d = OM_WAVE / 2; d > 0; d >>= 1) v += __shfl_xor(v, d, OM_WAVE);
    return v;
}

__device__ inline int wave_max(int v) {
#pragma unroll
    for (int d = OM_WAVE / 2; d > 0; d >>= 1) v = max(v, __shfl_xor(v, d, OM_WAVE));
    return v;
}

// lowest lane whose predicate holds, or -1
__device__ inline int first_lane(bool pred) {
    const unsigned long long m = __ballot(pred);
    return m ? (int)__builtin_ctzll(m) : -1;
}

// Window order.  A window's cost grows with its nonzeros (collection and
// cut-off passes) and distinct columns (layout, pairing, grouping), up to a
// few ms for R-MAT hub windows against tens of us for most.  Workgroup b takes
// order[b]: the windows sorted heaviest first by floor(log2(nnz + 1)), so the
// hub windows start first and, consecutive in the order, go to different
// XCDs (the dispatcher sends workgroup b to XCD b % 8).  Taking window b
// instead left one XCD running R-MAT's hub windows (mostly multiples of 8)
// long after the others had finished.
constexpr int OM_CLASSES = 32;

__device__ inline int window_class(const int32_t* __restrict__ rowptr, int64_t rows, int64_t nnz, int64_t w) {
    const int64_t r0 = w * OM_ROWS, r1 = min<int64_t>(r0 + OM_ROWS, rows);
    const int64_t a = min<int64_t>(max<int64_t>(rowptr[r0], 0), nnz), b = min<int64_t>(max<int64_t>(rowptr[r1], 0), nnz);
    return __clz((int)(b > a ? b - a : 0) + 1);           // 1 (heaviest) .. 31 (empty or one nonzero)
}

// cnt[c] += windows of class c
__global__ __launch_bounds__(256) void k_window_classes(const int32_t* __restrict__ rowptr, int64_t rows, int64_t nnz,
                                                        int64_t windows, int32_t* __restrict__ cnt) {
    __shared__ int32_t c[OM_CLASSES];
    if (threadIdx.x < OM_CLASSES) c[threadIdx.x] = 0;
    __syncthreads();
    const int64_t w = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (w < windows) atomicAdd(&c[window_class(rowptr, rows, nnz, w)], 1);
    __syncthreads();
    if (threadIdx.x < OM_CLASSES && c[threadIdx.x]) atomicAdd(&cnt[threadIdx.x], c[threadIdx.x]);
}

// order[start(c) + k] = the k-th window of class c to arrive (fill[] counts arrivals)
__global__ __launch_bounds__(256) void k_window_order(const int32_t* __restrict__ rowptr, int64_t rows, int64_t nnz,
                                                      int64_t windows, const int32_t* __restrict__ cnt,
                                                      int32_t* __restrict__ fill, int32_t* __restrict__ order) {
    __shared__ int32_t start[OM_CLASSES], local[OM_CLASSES], base[OM_CLASSES];
    if (threadIdx.x == 0) {
        int32_t run = 0;
        for (int k = 0; k < OM_CLASSES; ++k) {
            start[k] = run;
            run += cnt[k];
        }
    }
    if (threadIdx.x < OM_CLASSES) local[threadIdx.x] = 0;
    __syncthreads();
    const int64_t w = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    int cls = 0, rank = 0;
    if (w < windows) {
        cls = window_class(rowptr, rows, nnz, w);
        rank = atomicAdd(&local[cls], 1);
    }
    __syncthreads();
    if (threadIdx.x < OM_CLASSES && local[threadIdx.x])
        base[threadIdx.x] = start[threadIdx.x] + atomicAdd(&fill[threadIdx.x], local[threadIdx.x]);
    __syncthreads();
    if (w < windows) order[base[cls] + rank] = (int32_t)w;
}

__device__ inline int dedup_slot(int c) { return (int)(((uint32_t)c * 2654435761u) >> 20); }   // 12 bits
__device__ inline int count_slot(int c) { return (int)(((uint32_t)c * 2654435761u) >> 19); }   // 13 bits

// Collects the window's distinct columns in [0, limit] into a table: with
// MASKS the collection table (ids and row masks, OM_DEDUP slots), else the
// counting table (ids only, OM_COUNT slots in the same memory).  Stops once
// `cap` are in.  Returns the count (>= cap if it stopped), on every lane.
template <bool MASKS>
__device__ int collect(OnlineLds& s, OnlineCtl& ctl, const int32_t* __restrict__ col, int win_rows, int limit,
                       int cap) {
    constexpr int SLOTS = MASKS ? OM_DEDUP : OM_COUNT;
    int32_t* ids = MASKS ? s.a.d.id : s.a.cnt;
    __syncthreads();                       // earlier users of the shared memory are done
    for (int t = lane_id(); t < SLOTS; t += OM_WAVE) {
        ids[t] = -1;
        if (MASKS) s.a.d.mask[t] = 0;
    }
    if (lane_id() == 0) {
        ctl.occ = 0;
        ctl.stop = 0;
    }
    __syncthreads();
    auto stopped = [&]() { return __hip_atomic_load(&ctl.stop, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP) != 0; };
    // one flat pass over the window's nonzeros, OM_BATCH loads in flight per lane;
    // an entry's row is the number of row starts at or below it
    int32_t rp[OM_ROWS + 1];
#pragma unroll
    for (int k = 0; k <= OM_ROWS; ++k) rp[k] = ctl.rp[k];
    const int e_end = max(rp[0], rp[win_rows]);
    for (int e0 = rp[0]; e0 < e_end && !stopped(); e0 += OM_BATCH * OM_WAVE) {
        int cs[OM_BATCH];
#pragma unroll
        for (int u = 0; u < OM_BATCH; ++u) {
            const int e = e0 + u * OM_WAVE + lane_id();
            cs[u] = e < e_end ? col[e] : -1;
        }
#pragma unroll
        for (int u = 0; u < OM_BATCH; ++u) {
            const int c = cs[u];
            if (c < 0 || c > limit) continue;
            const int e = e0 + u * OM_WAVE + lane_id();
            int r = 0;
#pragma unroll
            for (int k = 1; k < OM_ROWS; ++k) r += (k < win_rows && e >= rp[k]) ? 1 : 0;
            // at most cap + OM_BATCH * 64 - 1 <= SLOTS / 2 + 256 columns are ever placed: a free slot exists
            for (int slot = MASKS ? dedup_slot(c) : count_slot(c), probe = 0; probe < SLOTS;
                 ++probe, slot = (slot + 1) & (SLOTS - 1)) {
                const int prev = atomicCAS(&ids[slot], -1, c);
                if (prev == -1 || prev == c) {
                    if (MASKS) atomicOr(&s.a.d.mask[slot], 1u << r);
                    if (prev == -1 && atomicAdd(&ctl.occ, 1) + 1 >= cap)
                        __hip_atomic_store(&ctl.stop, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
                    break;
                }
            }
        }
    }
    __syncthreads();
    return ctl.occ;
}

__global__ __launch_bounds__(OM_WAVE) void k_match_online(const int32_t* __restrict__ rowptr,
                                                          const int32_t* __restrict__ col, int64_t rows,
                                                          int64_t nnz, int dense_threshold, int t_max,
                                                          const int32_t* __restrict__ order,
                                                          unsigned long long* __restrict__ out) {
    __shared__ OnlineLds s;
    __shared__ OnlineCtl ctl;
    const int lane = lane_id();
    const int64_t r0 = (int64_t)order[blockIdx.x] * OM_ROWS;
    const int win_rows = (int)min<int64_t>(OM_ROWS, rows - r0);

    // ---- 1. the window's columns
    if (lane <= OM_ROWS) {
        const int64_t r = r0 + min(lane, win_rows);
        ctl.rp[lane] = (int32_t)min<int64_t>(max<int64_t>(rowptr[r], 0), nnz);   // clamped: reads stay inside col
    }
    if (lane == 0) ctl.collide = 0;
    __syncthreads();
    // largest column id and any negative one, in one read of the window
    const int e0 = ctl.rp[0], e1 = max(e0, ctl.rp[win_rows]);
    int hi = -1, neg = 0;
    for (int b = e0; b < e1; b += OM_BATCH * OM_WAVE) {
        int cs[OM_BATCH];
#pragma unroll
        for (int u = 0; u < OM_BATCH; ++u) {
            const int e = b + u * OM_WAVE + lane;
            cs[u] = e < e1 ? col[e] : -1;
            neg |= (e < e1) & (cs[u] < 0);
        }
#pragma unroll
        for (int u = 0; u < OM_BATCH; ++u) hi = max(hi, cs[u]);
    }
    hi = wave_max(hi);
    neg = wave_max(neg);
    if (neg) {
        if (lane == 0) atomicAdd(&out[10], 1ull);
        return;
    }
    if (hi < 0) return;                                    // empty window
    int n = collect<true>(s, ctl, col, win_rows, hi, OM_SLOTS + 1);
    if (n > OM_SLOTS) {
        // more distinct columns than the reference's table holds: keep the
        // OM_SLOTS smallest column ids (what inserting in ascending order
        // keeps).  Bisect for a cut whose distinct count lies in
        // [OM_SLOTS, OM_COUNT / 2] -- the counting table then holds every
        // distinct id below it -- and take the OM_SLOTS-th smallest of those.
        // count(lo) < OM_SLOTS and count(hi) > OM_COUNT / 2 cannot hold with
        // hi = lo + 1 (one more id adds at most one column), so the loop ends.
        if (lane == 0) atomicAdd(&out[9], 1ull);
        constexpr int BAND = OM_COUNT / 2;
        int lo = -1, cut = hi, m = 0;                      // count(lo) < OM_SLOTS
        for (;;) {
            m = collect<false>(s, ctl, col, win_rows, cut, BAND + 1);
            if (m >= OM_SLOTS && m <= BAND) break;
            if (m < OM_SLOTS) lo = cut;
            else hi = cut;
            cut = lo + (hi - lo) / 2;
        }
        // the m ids -> ids[0, m), sorted ascending
        {
            constexpr int PER = OM_COUNT / OM_WAVE;
            const int base = lane * PER;
            int c = 0, total = 0;
#pragma unroll 4
            for (int t = 0; t < PER; ++t) c += (s.a.cnt[base + t] != -1) ? 1 : 0;
            int at = wave_exclusive_scan(c, &total);
#pragma unroll 4      // re-read rather than keep PER values live across the scan
            for (int t = 0; t < PER; ++t)
                if (s.a.cnt[base + t] != -1) s.b.ids[at++] = (uint32_t)s.a.cnt[base + t];
        }
        int P = 2;
        while (P < m) P <<= 1;
        for (int t = m + lane; t < P; t += OM_WAVE) s.b.ids[t] = 0xffffffffu;
        __syncthreads();
        for (int k = 2; k <= P; k <<= 1) {
            for (int j = k >> 1; j > 0; j >>= 1) {
                for (int t = lane; t < P / 2; t += OM_WAVE) {
                    const int i = (t / j) * 2 * j + (t % j), q = i + j;
                    const uint32_t x = s.b.ids[i], y = s.b.ids[q];
                    if ((x > y) == ((i & k) == 0)) {
                        s.b.ids[i] = y;
                        s.b.ids[q] = x;
                    }
                }
                __syncthreads();
            }
        }
        hi = (int)s.b.ids[OM_SLOTS - 1];
        n = collect<true>(s, ctl, col, win_rows, hi, OM_DEDUP + 1);   // exactly OM_SLOTS columns
    }

    // collected columns -> keys[0, n)
    {
        constexpr int PER = OM_DEDUP / OM_WAVE;
        const int base = lane * PER;
        int cnt = 0, total = 0;
#pragma unroll 4
        for (int t = 0; t < PER; ++t) cnt += (s.a.d.id[base + t] != -1) ? 1 : 0;
        int at = wave_exclusive_scan(cnt, &total);
#pragma unroll 4
        for (int t = 0; t < PER; ++t)
            if (s.a.d.id[base + t] != -1)
                s.b.keys[at++] = ((uint64_t)(uint32_t)s.a.d.id[base + t] << 32) | s.a.d.mask[base + t];
    }
    __syncthreads();

    // ---- layout in the reference's table
    for (int t = lane; t < OM_SLOTS; t += OM_WAVE) {
        s.a.t.home_cnt[t] = 0;
        s.a.t.owner[t] = OM_FREE;
    }
    __syncthreads();
    int clash = 0;
    for (int k = lane; k < n; k += OM_WAVE)
        clash |= atomicAdd(&s.a.t.home_cnt[(int)(s.b.keys[k] >> 32) & (OM_SLOTS - 1)], 1u) != 0;
    clash = wave_max(clash);
    if (!clash) {
        for (int k = lane; k < n; k += OM_WAVE) s.a.t.owner[(int)(s.b.keys[k] >> 32) & (OM_SLOTS - 1)] = (uint16_t)k;
    } else {
        // ascending column order: bitonic sort of the keys
        int P = 2;
        while (P < n) P <<= 1;
        for (int t = n + lane; t < P; t += OM_WAVE) s.b.keys[t] = ~0ull;
        __syncthreads();
        for (int k = 2; k <= P; k <<= 1) {
            for (int j = k >> 1; j > 0; j >>= 1) {
                for (int t = lane; t < P / 2; t += OM_WAVE) {
                    const int i = (t / j) * 2 * j + (t % j);
                    const int q = i + j;
                    const uint64_t x = s.b.keys[i], y = s.b.keys[q];
                    if ((x > y) == ((i & k) == 0)) {
                        s.b.keys[i] = y;
                        s.b.keys[q] = x;
                    }
                }
                __syncthreads();
            }
        }
        // each column, in ascending order, takes the first free slot from its
        // home on (cyclic).  The free slots are a bitmap in registers: lane L
        // (L < 32) holds slots 64L .. 64L + 63.  One ballot finds the first
        // lane at or after the home's word with a free slot there, the lowest
        // set bit of its word is the slot; with none, the search wraps to the
        // first free slot of the table (k < n <= OM_SLOTS columns are placed
        // when column k is, so one exists).
        uint64_t free_w = lane < OM_SLOTS / 64 ? ~0ull : 0ull;
        for (int k0 = 0; k0 < n; k0 += OM_WAVE) {
            const int home_l = k0 + lane < n ? (int)(s.b.keys[k0 + lane] >> 32) & (OM_SLOTS - 1) : 0;
            const int cnt = min(OM_WAVE, n - k0);
            for (int j = 0; j < cnt; ++j) {
                const int h = __builtin_amdgcn_readlane(home_l, j);
                const int w = h >> 6;
                uint64_t cand = lane > w ? free_w : (lane == w ? free_w & (~0ull << (h & 63)) : 0ull);
                unsigned long long m = __ballot(cand != 0);
                if (m == 0) {
                    cand = free_w;
                    m = __ballot(cand != 0);
                }
                const int f = (int)__builtin_ctzll(m);
                const uint32_t lo = __builtin_amdgcn_readlane((uint32_t)cand, f);
                const uint32_t hi = __builtin_amdgcn_readlane((uint32_t)(cand >> 32), f);
                const int bit = lo ? __builtin_ctz(lo) : 32 + __builtin_ctz(hi);
                if (lane == f) free_w &= ~(1ull << bit);
                if (lane == 0) s.a.t.owner[f * 64 + bit] = (uint16_t)(k0 + j);
            }
        }
    }
    __syncthreads();
    // row masks in table order -> cmask[0, n)
    {
        constexpr int PER = OM_SLOTS / OM_WAVE;
        const int base = lane * PER;
        int cnt = 0, total = 0;
#pragma unroll 4
        for (int t = 0; t < PER; ++t) cnt += (s.a.t.owner[base + t] != OM_FREE) ? 1 : 0;
        int at = wave_exclusive_scan(cnt, &total);
#pragma unroll 4
        for (int t = 0; t < PER; ++t) {
            const int k = s.a.t.owner[base + t];
            if (k != OM_FREE) s.a.t.cmask[at++] = (uint16_t)(s.b.keys[k] & 0xFFFFu);
        }
    }
    __syncthreads();
    const uint16_t* cmask = s.a.t.cmask;

    // ---- 2. 1:2 pairs
    // Column i's candidates i + 1 .. i + 32 sit on lanes 0 .. 31 (lane l holds
    // cmask[i + 1 + l], loaded one column ahead).  Which of the next 64
    // columns are already paired is a wave-uniform bit mask that shifts with
    // i (a partner lies at most 32 ahead), so no step waits on shared memory.
    int np = 0, nd = 0, fine = 0;
    uint64_t taken = 0;                                    // bit d: column i + d is paired
    uint32_t cand = cmask[min(1 + lane, n - 1)];
    uint32_t mi = cmask[0];
    for (int i = 0; i < n; ++i) {
        const uint32_t cand_next = cmask[min(i + 2 + lane, n - 1)];
        if (!(taken & 1)) {
            const bool valid = (lane < OM_PAIR_SEARCH) & (i + 1 + lane < n) & !((taken >> (1 + (lane & 31))) & 1);
            const int cost = valid ? __popc(mi & cand) : 64;
            int best = -1;
            for (int c = 0; c <= min(t_max, OM_ROWS); ++c) {
                const int l = first_lane(cost == c);
                if (l >= 0) {
                    best = l;
                    break;
                }
            }
            if (best >= 0) {
                const uint32_t mb = __builtin_amdgcn_readlane(cand, best);
                if (lane == 0) {
                    s.b.p.pk[np] = (mi & mb) | ((mi | mb) << 16);
                    s.b.p.c1[np] = (uint16_t)i;
                    s.b.p.c2[np] = (uint16_t)(i + 1 + best);
                }
                taken |= 2ull << best;
                ++np;
            } else {
                if (lane == 0) s.b.p.pool[nd] = (uint16_t)i;
                ++nd;
                ++fine;
            }
        }
        taken >>= 1;
        mi = __builtin_amdgcn_readlane(cand, 0);           // cmask[i + 1]
        cand = cand_next;
    }
    __syncthreads();

    // ---- 3. 2:4 groups
    // A grouped pair's word is set to all ones: its row union (high half) is
    // never empty, so it never passes the test again.  One wave: lane 0's
    // writes are ordered before the next pair's reads by a wave-scope fence
    // (shared memory serves one wave's accesses in order).  A pair's nonzeros
    // are popc(both) + popc(either).
    int groups = 0, coarse = 0, fake = 0;
    for (int i = 0; i < np; ++i) {
        const uint32_t me = s.b.p.pk[i];
        if (me == ~0u) continue;
        const uint32_t a1 = me & 0xFFFFu, a2 = me >> 16;
        int best = -1;
        uint32_t other = 0;
        for (int step = 0; step < OM_GROUP_SEARCH; step += OM_WAVE) {
            const int j = i + 1 + step + lane;
            const uint32_t o = j < np ? s.b.p.pk[j] : ~0u;
            const int l = first_lane(((a1 & (o >> 16)) | (a2 & o)) == 0);
            if (l >= 0) {
                best = i + 1 + step + l;
                other = __builtin_amdgcn_readlane(o, l);
                break;
            }
            if (i + 1 + step + OM_WAVE >= np) break;
        }
        if (best >= 0) {
            ++groups;
            fake += 32 - (__popc(me) + __popc(other));
        } else {
            coarse += 2;
        }
        if (lane == 0) {
            if (best >= 0) {
                s.b.p.pk[best] = ~0u;
            } else {
                s.b.p.pool[nd] = s.b.p.c1[i];
                s.b.p.pool[nd + 1] = s.b.p.c2[i];
            }
        }
        if (best < 0) nd += 2;
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    }
    __syncthreads();

    // ---- 4. the pool
    const int tc = nd / dense_threshold;
    const int processed = tc * dense_threshold;
    const int residue = nd - processed;
    const int rgroups = (residue + 1) / 2;
    int rnnz = 0;
    for (int k = processed + lane; k < nd; k += OM_WAVE) rnnz += __popc(cmask[s.b.p.pool[k]]);
    rnnz = wave_sum(rnnz);
    fake += 32 * rgroups - rnnz;
    const int sptc = groups + rgroups;
    const int blocks = (sptc + 3) / 4;

    // ---- 5. totals
    if (lane == 0) {
        const unsigned long long v[9] = {(unsigned long long)sptc, (unsigned long long)tc,
                                         (unsigned long long)processed, (unsigned long long)residue,
                                         (unsigned long long)fine, (unsigned long long)coarse,
                                         (unsigned long long)blocks, (unsigned long long)(blocks * 4 - sptc),
                                         (unsigned long long)fake};
        for (int k = 0; k < 9; ++k)
            if (v[k]) atomicAdd(&out[k], v[k]);
    }
}

__global__ void k_zero_u64(unsigned long long* p, int n) {
    const int i = (int)(blockIdx.x * blockDim.x + threadIdx.x);
    if (i < n) p[i] = 0;
}

}  // namespace

// the statistics of one call: counters zeroed, windows ordered heaviest first
// into `scratch` (2 * OM_CLASSES + windows int32: class counts, class fill
// counters, the window order; cleared here by a kernel, not a memset), then
// one workgroup per window
static int match_online_launch(const int32_t* d_rowptr, const int32_t* d_col, int64_t rows, int64_t nnz,
                               int dense_threshold, int t_max, unsigned long long* d_out, int32_t* scratch,
                               int64_t windows, hipStream_t stream) {
    int32_t* cnt = scratch;
    int32_t* fill = scratch + OM_CLASSES;
    int32_t* order = scratch + 2 * OM_CLASSES;
    hipLaunchKernelGGL(k_zero_u64, dim3(1), dim3(64), 0, stream, d_out, TRSPMM_ONLINE_COUNTERS);
    hipLaunchKernelGGL(k_zero_u64, dim3(1), dim3(64), 0, stream, (unsigned long long*)scratch, OM_CLASSES);
    const unsigned blocks = (unsigned)((windows + 255) / 256);
    hipLaunchKernelGGL(k_window_classes, dim3(blocks), dim3(256), 0, stream, d_rowptr, rows, nnz, windows, cnt);
    hipLaunchKernelGGL(k_window_order, dim3(blocks), dim3(256), 0, stream, d_rowptr, rows, nnz, windows, cnt, fill,
                       order);
    hipLaunchKernelGGL(k_match_online, dim3((unsigned)windows), dim3(OM_WAVE), 0, stream, d_rowptr, d_col, rows, nnz,
                       dense_threshold, t_max, order, d_out);
    return hipGetLastError() == hipSuccess ? TRSPMM_OK : TRSPMM_EHIP;
}

static int check_online_args(int64_t rows, int64_t nnz, int dense_threshold) {
    TRSPMM_CHECK(rows > 0 && rows < (int64_t)1 << 31, "match_2to4_online: rows must be in [1, 2^31)");
    TRSPMM_CHECK(nnz >= 0 && nnz < (int64_t)1 << 31, "match_2to4_online: nnz must be in [0, 2^31)");
    TRSPMM_CHECK(dense_threshold >= 1, "match_2to4_online: dense_threshold must be >= 1");
    return TRSPMM_OK;
}

int match_online_device(const int32_t* d_rowptr, const int32_t* d_col, int64_t rows, int64_t nnz,
                        int dense_threshold, int t_max, unsigned long long* d_out, hipStream_t stream) {
    int rc;
    if ((rc = check_online_args(rows, nnz, dense_threshold))) return rc;
    const int64_t windows = (rows + OM_ROWS - 1) / OM_ROWS;
    // stream-ordered scratch: class counts, class fill counters, the window order
    int32_t* scratch = nullptr;
    TRSPMM_HIP(hipMallocAsync((void**)&scratch, 4 * (2 * OM_CLASSES + windows), stream));
    rc = match_online_launch(d_rowptr, d_col, rows, nnz, dense_threshold, t_max, d_out, scratch, windows, stream);
    TRSPMM_HIP(hipFreeAsync(scratch, stream));
    if (rc) set_error("match_2to4_online: kernel launch failed");
    return rc;
}

// ---------------------------------------------------------------- allocation-free variant
//
// Scratch of the allocation-free entry point: file-scope device memory, one
// copy per GPU, held until the process ends.  Host code never passes the
// array's name to a kernel (on the host it is a stand-in, not a device
// address): each device's address is taken once, with that device current,
// by hipGetSymbolAddress, and that address is what the kernels receive.
// Every call on a GPU shares its copy, which keeps the last call's contents:
// the counters a call needs at zero are cleared by a kernel at its start, and
// calls that overlap on one GPU (concurrent streams) race on it.
constexpr int64_t kOnlineScratchInts = (int64_t)1 << 21;       // 8 MiB
constexpr int64_t kOnlineMaxWindows = kOnlineScratchInts - 2 * OM_CLASSES;
static __device__ __align__(256) int32_t scratch[kOnlineScratchInts];

namespace {
constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
int32_t* g_scratch_dev[kMaxDevices] = {};

// the current device's address of `scratch`, looked up on the device's first call
int online_scratch(int32_t** out) {
    int dev = 0;
    TRSPMM_HIP(hipGetDevice(&dev));
    TRSPMM_CHECK(dev >= 0 && dev < kMaxDevices, "match_2to4_online_noalloc: device index out of range");
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (g_scratch_dev[dev] == nullptr) {
        void* p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || p == nullptr) {
            set_error(std::string("match_2to4_online_noalloc: hipGetSymbolAddress(scratch) failed: ") +
                      hipGetErrorString(e));
            return TRSPMM_EHIP;
        }
        g_scratch_dev[dev] = (int32_t*)p;
    }
    *out = g_scratch_dev[dev];
    return TRSPMM_OK;
}
}  // namespace

int match_online_device_noalloc(const int32_t* d_rowptr, const int32_t* d_col, int64_t rows, int64_t nnz,
                                int dense_threshold, int t_max, unsigned long long* d_out, hipStream_t stream) {
    int rc;
    if ((rc = check_online_args(rows, nnz, dense_threshold))) return rc;
    const int64_t windows = (rows + OM_ROWS - 1) / OM_ROWS;
    TRSPMM_CHECK(windows <= kOnlineMaxWindows,
                 "match_2to4_online_noalloc: at most " + std::to_string(kOnlineMaxWindows * OM_ROWS) +
                     " rows (its fixed scratch holds the order of " + std::to_string(kOnlineMaxWindows) +
                     " windows); use trspmm_match_2to4_online");
    int32_t* s = nullptr;
    if ((rc = online_scratch(&s))) return rc;
    rc = match_online_launch(d_rowptr, d_col, rows, nnz, dense_threshold, t_max, d_out, s, windows, stream);
    if (rc) set_error("match_2to4_online_noalloc: kernel launch failed");
    return rc;
}

}  // namespace trspmm
