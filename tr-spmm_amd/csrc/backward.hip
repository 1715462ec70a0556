// Backward-pass building blocks of the routed SpMM (no counterpart in the
// reference, which has no backward):
//
//   csr_transpose  A (rows x ncols, canonical CSR) -> A^T as a canonical CSR
//                  plus the edge permutation perm[i] = A-edge of A^T entry i.
//                  A stable radix sort of the edges by column keeps the rows of
//                  each column ascending, so A^T is canonical and its routed
//                  plan is the reference router's tiling of A^T.
//   sddmm          dvals[e] = sum_f X[row_e, f] * Y[col_e, f] over A's pattern
//                  (X = dY rows x K, Y = B ncols x K, fp16/bf16, fp32 result):
//                  the gradient of sum(C * dY), C = A·B, with respect to A's
//                  values, in A's edge order.
#include <cstdlib>
#include <cstring>   // before rocPRIM: its host code calls memset

#include <rocprim/rocprim.hpp>

#include <type_traits>
#include <vector>

#include "backward.h"
#include "router.h"
#include "scan.h"

namespace trspmm {

namespace {

inline int nblk(int64_t n, int per = 256) { return (int)((n + per - 1) / per); }

__global__ void k_check_cols(const int32_t* __restrict__ col, int64_t nnz, int64_t ncols, int32_t* __restrict__ err) {
    const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e < nnz && (col[e] < 0 || col[e] >= ncols)) atomicOr(err, 1);
}

__global__ void k_iota(int32_t* __restrict__ v, int64_t n) {
    const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e < n) v[e] = (int32_t)e;
}

// row of edge e: the last r with rowptr[r] <= e (rowptr non-decreasing)
__device__ inline int32_t row_of(const int32_t* __restrict__ rowptr, int64_t rows, int32_t e) {
    int64_t lo = 0, hi = rows;            // invariant: rowptr[lo] <= e < rowptr[hi]
    while (hi - lo > 1) {
        const int64_t mid = (lo + hi) >> 1;
        if (rowptr[mid] <= e) lo = mid;
        else hi = mid;
    }
    return (int32_t)lo;
}

__global__ void k_transpose_fill(const int32_t* __restrict__ rowptr, int64_t rows, const float* __restrict__ val,
                                 const int32_t* __restrict__ perm, int64_t nnz, int32_t* __restrict__ t_col,
                                 float* __restrict__ t_val) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= nnz) return;
    const int32_t e = perm[i];
    t_col[i] = row_of(rowptr, rows, e);
    if (t_val) t_val[i] = val[e];
}

// t_rowptr[c] = first sorted position with column >= c, c = 0..ncols
__global__ void k_transpose_rowptr(const int32_t* __restrict__ skeys, int64_t nnz, int64_t ncols,
                                   int32_t* __restrict__ t_rowptr) {
    const int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (c > ncols) return;
    t_rowptr[c] = (int32_t)lower_bound<int32_t>(skeys, nnz, (int32_t)c);
}

// ------------------------------------------------------------------ SDDMM
//
// Lane groups of LPR lanes; a group covers 8*LPR features of a row with
// 16-byte loads (more passes when K is wider) and takes U = 8 consecutive
// edges per batch from its own contiguous range of EPG edges.  Each lane
// accumulates its features' partial dot product for the 8 edges; a
// reduce-scatter butterfly over the group (halving the edge set at each
// lane-bit) leaves every edge's full sum in 8/LPR lane(s), about one shuffle
// per edge.
constexpr int SDDMM_U = 8;
constexpr int SDDMM_EPG = 256;           // edges per lane group
constexpr int SDDMM_BLOCK = 256;

// 16-bit value in the low half of `w` as fp32
template <typename DT>
__device__ inline float lo16(uint32_t w) {
    if constexpr (DT::code == TRSPMM_BF16) return __builtin_bit_cast(float, w << 16);
    else return (float)__builtin_bit_cast(_Float16, (uint16_t)(w & 0xffffu));
}

// dot product of 8 packed 16-bit values.  The halves are taken from the 32-bit
// words with shifts: bit-casting the load (or one of its 32-bit words) to a
// vector of 16-bit values and using its elements compiled to repeated reads
// of element 0 (hipcc, ROCm 7.2, gfx950), with fp32 conversions and with the
// packed dot instructions alike.
template <typename DT>
__device__ inline float dot8(u32x4 a, u32x4 b) {
    float s = 0.f;
#pragma unroll
    for (int i = 0; i < 4; ++i) {
        s = __builtin_fmaf(lo16<DT>(a[i]), lo16<DT>(b[i]), s);
        s = __builtin_fmaf(lo16<DT>(a[i] >> 16), lo16<DT>(b[i] >> 16), s);
    }
    return s;
}

// one butterfly step over H pairs: lanes with `hi` keep part[H..2H), the others part[0..H)
template <int H>
__device__ inline void rs_step(float (&part)[SDDMM_U], bool hi, int bit) {
#pragma unroll
    for (int k = 0; k < H; ++k) {
        const float send = hi ? part[k] : part[k + H];
        const float keep = hi ? part[k + H] : part[k];
        part[k] = keep + __shfl_xor(send, bit, 64);
    }
}

// Reduce-scatter of NE (2, 4 or 8) per-edge partial sums (part[0..NE)) over a
// lane group of LPR lanes (j = lane within the group): at each lane-bit keep
// the half of the edge set that bit selects and add the partner lane's copy
// of it; lane-bits left once each lane holds one edge are plain all-reduces.
// Afterwards the lane holds the full sums of edges [base, base + NV) in
// part[0..NV), and the lanes whose bits in LOWMASK are zero are the writers.
template <int NE, int BIT>
__device__ inline void rs_steps(float (&part)[SDDMM_U], int j, int& base) {
    if constexpr (BIT >= 1) {
        if constexpr (NE > 1) {
            const bool hi = (j & BIT) != 0;
            rs_step<NE / 2>(part, hi, BIT);
            base += hi ? NE / 2 : 0;
            rs_steps<NE / 2, BIT / 2>(part, j, base);
        } else {
            part[0] += __shfl_xor(part[0], BIT, 64);
            rs_steps<1, BIT / 2>(part, j, base);
        }
    }
}

template <int LPR, int NE = SDDMM_U>
struct EdgeReduce {
    static_assert(NE >= 1 && NE <= SDDMM_U && (NE & (NE - 1)) == 0, "NE must be a power of two <= 8");
    static constexpr int NV = NE >= LPR ? NE / LPR : 1;
    static constexpr int LOWMASK = NE >= LPR ? 0 : LPR / NE - 1;
    __device__ static inline int run(float (&part)[SDDMM_U], int j) {
        int base = 0;
        rs_steps<NE, LPR / 2>(part, j, base);
        return base;
    }
};

template <int LPR, typename DT>
__global__ __launch_bounds__(SDDMM_BLOCK) void k_sddmm(const int32_t* __restrict__ rowptr, int64_t rows,
                                                       const int32_t* __restrict__ col, int64_t nnz,
                                                       const uint16_t* __restrict__ X, int64_t ldx,
                                                       const uint16_t* __restrict__ Y, int64_t ldy, int K,
                                                       int64_t ncols, float* __restrict__ out) {
    constexpr int U = SDDMM_U;
    const int lane = threadIdx.x & 63;
    const int j = lane % LPR;
    const int64_t group = ((int64_t)blockIdx.x * SDDMM_BLOCK + threadIdx.x) / LPR;
    const int64_t g0 = group * SDDMM_EPG;
    if (g0 >= nnz) return;                // the group leaves together; no barriers below
    const int64_t g1 = g0 + SDDMM_EPG < nnz ? g0 + SDDMM_EPG : nnz;
    const int64_t elast = nnz - 1;
    int32_t row = row_of(rowptr, rows, (int32_t)g0);
    int32_t row_end = rowptr[row + 1];

    for (int64_t eb = g0; eb < g1; eb += U) {
        int32_t c[U], r[U];
        bool ok[U];
#pragma unroll
        for (int u = 0; u < U; ++u) {
            const int64_t e = eb + u < elast ? eb + u : elast;   // clamped: unconditional loads
            const int32_t cu = col[e];
            ok[u] = cu >= 0 && (int64_t)cu < ncols;      // out-of-range column: reads row 0, gives 0
            c[u] = ok[u] ? cu : 0;
            if (eb + u < g1) {
                while (e >= row_end) {
                    ++row;
                    row_end = rowptr[row + 1];
                }
            }
            r[u] = row;
        }
        float part[U];
#pragma unroll
        for (int u = 0; u < U; ++u) part[u] = 0.f;
        for (int f = 8 * j; f < K; f += 8 * LPR) {
            u32x4 xv[U], yv[U];
#pragma unroll
            for (int u = 0; u < U; ++u) {
                xv[u] = *(const u32x4*)(X + (int64_t)r[u] * ldx + f);
                yv[u] = *(const u32x4*)(Y + (int64_t)c[u] * ldy + f);
            }
#pragma unroll
            for (int u = 0; u < U; ++u) part[u] += dot8<DT>(xv[u], yv[u]);
        }
#pragma unroll
        for (int u = 0; u < U; ++u) part[u] = ok[u] ? part[u] : 0.f;
        using R = EdgeReduce<LPR>;
        const int base = R::run(part, j);
        if ((j & R::LOWMASK) == 0) {
#pragma unroll
            for (int k = 0; k < R::NV; ++k) {
                const int64_t e = eb + base + k;
                if (e < g1) out[e] = part[k];
            }
        }
    }
}

template <typename DT>
int launch_sddmm_typed(const int32_t* rowptr, int64_t rows, const int32_t* col, int64_t nnz, int64_t ncols,
                       const void* X, int64_t ldx, const void* Y, int64_t ldy, int64_t K, float* out, hipStream_t s) {
    const int lpr = K >= 128 ? 16 : (K >= 64 ? 8 : 4);
    const int64_t groups = (nnz + SDDMM_EPG - 1) / SDDMM_EPG;
    const int64_t threads = groups * lpr;
    const dim3 grid((unsigned)((threads + SDDMM_BLOCK - 1) / SDDMM_BLOCK));
    const auto* x = (const uint16_t*)X;
    const auto* y = (const uint16_t*)Y;
    if (lpr == 16)
        hipLaunchKernelGGL((k_sddmm<16, DT>), grid, dim3(SDDMM_BLOCK), 0, s, rowptr, rows, col, nnz, x, ldx, y, ldy,
                           (int)K, ncols, out);
    else if (lpr == 8)
        hipLaunchKernelGGL((k_sddmm<8, DT>), grid, dim3(SDDMM_BLOCK), 0, s, rowptr, rows, col, nnz, x, ldx, y, ldy,
                           (int)K, ncols, out);
    else
        hipLaunchKernelGGL((k_sddmm<4, DT>), grid, dim3(SDDMM_BLOCK), 0, s, rowptr, rows, col, nnz, x, ldx, y, ldy,
                           (int)K, ncols, out);
    TRSPMM_HIP(hipGetLastError());
    return TRSPMM_OK;
}


// ------------------------------------------------------------------ fused A^T backward
//
// One pass over the rows of A^T (row c of A^T = column c of A).  For an entry
// (c, r, v) of A^T the dY row r is gathered once and used twice:
//     dB[c]    += v * dY[r]           (dB = A^T dY, fp32 accumulation)
//     dvals_t  =  <dY[r], B[c]>       (the SDDMM value gradient, A^T order)
// A lane group of LPR lanes owns whole rows: it keeps B[c] (its features,
// NCH 16-byte chunks per lane) and the dB accumulator in registers, streams
// the row's entries U = 8 at a time (U * NCH loads per lane in flight) and
// writes the dB row once.  Work items (built on the host from A^T's row
// pointers): kind 0 = a run of whole rows [r0, r1) shared round-robin by the
// wave's groups (plain stores); kind 1 = a chunk [b, e) of one long row split
// over the groups (fp32 atomic adds into a row zeroed beforehand).
constexpr int BWD_BLOCK = 256;
constexpr int BWD_U = 4;                // entries per batch (two batches in flight)
// Item shape, from a sweep on R-MAT scale 18 at K = 256 (fused backward 0.85 -> 0.76 ms):
// split only rows above 1024 entries (each split row costs fp32 atomics), and
// keep kind-0 items small (<= 512 entries, <= 8 rows) so the tail balances.
constexpr int BWD_ROW_CAP = 1024;       // entries of one row handled as a whole
constexpr int BWD_ITEM_CAP = 512;       // entries of a kind-0 item
constexpr int BWD_ITEM_ROWS = 8;        // rows of a kind-0 item

template <typename DT>
__device__ inline void fma8(float (&acc)[8], float v, u32x4 x) {
#pragma unroll
    for (int i = 0; i < 4; ++i) {
        acc[2 * i] = __builtin_fmaf(v, lo16<DT>(x[i]), acc[2 * i]);
        acc[2 * i + 1] = __builtin_fmaf(v, lo16<DT>(x[i] >> 16), acc[2 * i + 1]);
    }
}

template <int LPR, int NCH, typename DT>
__device__ inline void bwd_row_range(int64_t c, int32_t e0, int32_t e1, bool atomic, int j,
                                     const int32_t* __restrict__ col, const uint16_t* __restrict__ val,
                                     const uint16_t* __restrict__ dY, int64_t lddy,
                                     const uint16_t* __restrict__ B, int64_t ldb, int K,
                                     float* __restrict__ dB, int64_t lddb, float* __restrict__ dvt) {
    constexpr int U = BWD_U;
    using R = EdgeReduce<LPR, U>;
    u32x4 bc[NCH];
    float acc[NCH][8];
#pragma unroll
    for (int k = 0; k < NCH; ++k) {
        const int f = 8 * j + 8 * LPR * k;
        bc[k] = f < K ? *(const u32x4*)(B + c * ldb + f) : u32x4{0u, 0u, 0u, 0u};
#pragma unroll
        for (int q = 0; q < 8; ++q) acc[k][q] = 0.f;
    }
    // Batches of U entries in a two-ahead pipeline, unrolled by two so the ring
    // slots are compile-time: batch b's dY rows were gathered during batch
    // b - 1 from indices loaded during batch b - 2, so computing b waits only
    // for loads issued before the current batch's.  Every load is unconditional
    // (indices clamped to the last entry; values masked where they are used), so
    // the count of loads in flight is the same on every path.  An odd batch
    // count gets one masked extra batch.
    const int nb = (e1 - e0 + U - 1) / U;
    // A batch's U column indices and values are loaded once per lane group:
    // lane j loads entry j % U, and the group takes them with lane shuffles
    // (every lane loading every entry returned each index LPR times, and the
    // vector-memory data returned is what such gathers are bound by).
    const int gbase = (int)(threadIdx.x & 63) - j;
    int32_t ridx[2];
    uint32_t vraw[2];
    u32x4 y[2][U][NCH];
    auto load_idx = [&](int b, int32_t& r) {
        const int32_t e = e0 + b * U + (j & (U - 1));
        r = col[e < e1 ? e : e1 - 1];
    };
    auto load_val = [&](int b, uint32_t& v) {
        const int32_t e = e0 + b * U + (j & (U - 1));
        v = val[e < e1 ? e : e1 - 1];
    };
    auto gather = [&](int32_t r_lane, u32x4 (&yy)[U][NCH]) {
#pragma unroll
        for (int u = 0; u < U; ++u) {
            const int32_t r = __shfl(r_lane, gbase + u, 64);
#pragma unroll
            for (int k = 0; k < NCH; ++k) {
                const int f = 8 * j + 8 * LPR * k;
                yy[u][k] = *(const u32x4*)(dY + (int64_t)r * lddy + (f < K ? f : 0));
            }
        }
    };
    auto compute = [&](int b, const u32x4 (&yy)[U][NCH], uint32_t v_lane) {
        const int32_t eb = e0 + b * U;
        float part[SDDMM_U];
#pragma unroll
        for (int u = 0; u < U; ++u) {
            const uint32_t vb = (uint32_t)__shfl((int)v_lane, gbase + u, 64);
            const float vu = eb + u < e1 ? lo16<DT>(vb) : 0.f;
            part[u] = 0.f;
#pragma unroll
            for (int k = 0; k < NCH; ++k) {
                if (8 * j + 8 * LPR * k < K) {
                    fma8<DT>(acc[k], vu, yy[u][k]);
                    part[u] += dot8<DT>(yy[u][k], bc[k]);
                }
            }
        }
        const int base = R::run(part, j);
        if ((j & R::LOWMASK) == 0) {
#pragma unroll
            for (int k = 0; k < R::NV; ++k)
                if (eb + base + k < e1) dvt[eb + base + k] = part[k];
        }
    };
    if (nb > 0) {
        load_idx(0, ridx[0]);
        load_idx(1, ridx[1]);
        gather(ridx[0], y[0]);
        load_val(0, vraw[0]);
        auto body = [&](int b, auto par_tag) {
            constexpr int PAR = decltype(par_tag)::value;
            load_idx(b + 2, ridx[PAR]);              // batch b + 2
            gather(ridx[1 - PAR], y[1 - PAR]);       // batch b + 1
            load_val(b + 1, vraw[1 - PAR]);
            compute(b, y[PAR], vraw[PAR]);           // batch b
        };
        for (int b = 0; b < nb; b += 2) {
            body(b, std::integral_constant<int, 0>{});
            body(b + 1, std::integral_constant<int, 1>{});
        }
    }
#pragma unroll
    for (int k = 0; k < NCH; ++k) {
        const int f = 8 * j + 8 * LPR * k;
        if (f >= K) continue;
        float* dst = dB + c * lddb + f;
        if (!atomic) {
            *(f32x4*)dst = f32x4{acc[k][0], acc[k][1], acc[k][2], acc[k][3]};
            *(f32x4*)(dst + 4) = f32x4{acc[k][4], acc[k][5], acc[k][6], acc[k][7]};
        } else {
#pragma unroll
            for (int q = 0; q < 8; ++q)
                if (acc[k][q] != 0.f) unsafeAtomicAdd(dst + q, acc[k][q]);
        }
    }
}

template <int LPR, int NCH, typename DT>
__global__ __launch_bounds__(BWD_BLOCK) void k_bwd_fused(const int4* __restrict__ items, int64_t n_items,
                                                        const int32_t* __restrict__ rowptr,
                                                        const int32_t* __restrict__ col,
                                                        const uint16_t* __restrict__ val,
                                                        const uint16_t* __restrict__ dY, int64_t lddy,
                                                        const uint16_t* __restrict__ B, int64_t ldb, int K,
                                                        float* __restrict__ dB, int64_t lddb,
                                                        float* __restrict__ dvt) {
    constexpr int G = 64 / LPR;
    const int64_t it = (int64_t)blockIdx.x * (BWD_BLOCK / 64) + threadIdx.x / 64;
    if (it >= n_items) return;                    // whole waves leave; no block barriers below
    const int lane = threadIdx.x & 63, g = lane / LPR, j = lane % LPR;
    const int4 item = items[it];
    if (item.x == 0) {
        for (int32_t c = item.y + g; c < item.z; c += G)
            bwd_row_range<LPR, NCH, DT>(c, rowptr[c], rowptr[c + 1], false, j, col, val, dY, lddy, B, ldb, K, dB,
                                        lddb, dvt);
    } else {
        const int32_t n = item.w - item.z, per = (n + G - 1) / G;
        const int32_t b = item.z + g * per, e = b + per < item.w ? b + per : item.w;
        if (b < e)
            bwd_row_range<LPR, NCH, DT>(item.y, b, e, true, j, col, val, dY, lddy, B, ldb, K, dB, lddb, dvt);
    }
}

__global__ void k_zero_list(const int32_t* __restrict__ rows, int64_t n, int K, float* __restrict__ dst,
                            int64_t ld) {
    const int64_t i = blockIdx.x;
    if (i >= n) return;
    for (int f = threadIdx.x; f < K; f += blockDim.x) dst[(int64_t)rows[i] * ld + f] = 0.f;
}

__global__ void k_invert(const int32_t* __restrict__ perm, int64_t n, int32_t* __restrict__ inv) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) inv[perm[i]] = (int32_t)i;
}

// dst[i] = src[idx[i]].  A block of 256 lanes covers a span of
// kGatherSpan entries, each lane GATHER_PER of them 64 apart (every load and
// store instruction coalesced, GATHER_PER random reads in flight per lane).
// The grid is 8 * per blocks and block b takes span (b % 8) * per + b / 8, so
// XCD x (blocks go to the 8 XCDs round robin) walks one contiguous eighth of
// dst: entries of nearby rows of A that share a column sit next to each other
// in src, and their lines then meet in one L2.  (Fused backward on R-MAT scale
// 18, K = 256, in one process: 1 or 2 entries per lane 434.5-435.3 us, 4 or 8
// 437.0-437.7 us, 16 443.8-444.9 us.)
constexpr int GATHER_PER = 2;
constexpr int64_t kGatherSpan = 256 * GATHER_PER;

__global__ __launch_bounds__(256) void k_gather_f32(const float* __restrict__ src, const int32_t* __restrict__ idx,
                                                    int64_t n, int64_t per, float* __restrict__ dst) {
    const int64_t span = (int64_t)(blockIdx.x % 8) * per + blockIdx.x / 8;
    const int64_t i0 = span * kGatherSpan + (threadIdx.x / 64) * (64 * GATHER_PER) + (threadIdx.x & 63);
    float v[GATHER_PER];
#pragma unroll
    for (int k = 0; k < GATHER_PER; ++k) {
        const int64_t i = i0 + 64 * k;
        v[k] = i < n ? src[idx[i]] : 0.f;
    }
#pragma unroll
    for (int k = 0; k < GATHER_PER; ++k)
        if (i0 + 64 * k < n) dst[i0 + 64 * k] = v[k];
}

__global__ void k_to16(const float* __restrict__ src, int64_t n, int dtype, uint16_t* __restrict__ dst) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) dst[i] = dtype == TRSPMM_BF16 ? f32_to_bf16_bits(src[i]) : f32_to_f16_rne(src[i]);
}

// dst[inv[e]] = src[e] converted as k_to16
__global__ void k_scatter16(const float* __restrict__ src, const int32_t* __restrict__ inv, int64_t n, int dtype,
                            uint16_t* __restrict__ dst) {
    const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e < n) dst[inv[e]] = dtype == TRSPMM_BF16 ? f32_to_bf16_bits(src[e]) : f32_to_f16_rne(src[e]);
}

// pos[perm[e]] = base + i for the A^T edge e of slot i (a record, or a TC
// value): the slot of D that holds each A edge's value gradient
__global__ void k_pos_from_records(const int32_t* __restrict__ edge, int64_t n, const int32_t* __restrict__ perm,
                                   int64_t nnz, int64_t base, int32_t* __restrict__ pos) {
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
        const int32_t e = edge[i];
        if (e >= 0 && e < nnz) pos[perm[e]] = (int32_t)(base + i);
    }
}

__global__ void k_count_negative(const int32_t* __restrict__ v, int64_t n, int32_t* __restrict__ cnt) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n && v[i] < 0) atomicAdd(cnt, 1);
}

// err |= 1 unless inv[perm[i]] == i for every i (perm entries already in range)
__global__ void k_check_inverse(const int32_t* __restrict__ perm, const int32_t* __restrict__ inv, int64_t n,
                                int32_t* __restrict__ err) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n && inv[perm[i]] != (int32_t)i) atomicOr(err, 1);
}

template <typename DT>
int launch_bwd_typed(const BwdPlan& P, const void* dY, int64_t lddy, const void* B, int64_t ldb, int64_t K,
                     float* dB, int64_t lddb, float* dvals, hipStream_t s) {
    if (P.n_zero > 0) {
        hipLaunchKernelGGL(k_zero_list, dim3((unsigned)P.n_zero), dim3(256), 0, s, P.zero_rows.as<int32_t>(),
                           P.n_zero, (int)K, dB, lddb);
        TRSPMM_HIP(hipGetLastError());
    }
    const dim3 grid((unsigned)((P.n_items + BWD_BLOCK / 64 - 1) / (BWD_BLOCK / 64)));
    const auto* y = (const uint16_t*)dY;
    const auto* b = (const uint16_t*)B;
    const auto* items = P.items.as<int4>();
    const auto* rp = P.rowptr.as<int32_t>();
    const auto* cl = P.col.as<int32_t>();
    const auto* vl = P.val.as<uint16_t>();
    float* dvt = P.dvals_t.as<float>();
#define TRSPMM_BWD(LPR, NCH)                                                                                      \
    hipLaunchKernelGGL((k_bwd_fused<LPR, NCH, DT>), grid, dim3(BWD_BLOCK), 0, s, items, P.n_items, rp, cl, vl, y,  \
                       lddy, b, ldb, (int)K, dB, lddb, dvt)
    if (P.n_items > 0) {
        if (K <= 64) TRSPMM_BWD(8, 1);
        else if (K <= 128) TRSPMM_BWD(16, 1);
        else if (K <= 256) TRSPMM_BWD(32, 1);
        else TRSPMM_BWD(32, 2);
        TRSPMM_HIP(hipGetLastError());
    }
#undef TRSPMM_BWD
    if (P.nnz > 0) {
        const int64_t per = ((P.nnz + kGatherSpan - 1) / kGatherSpan + 7) / 8;
        hipLaunchKernelGGL(k_gather_f32, dim3((unsigned)(8 * per)), dim3(256), 0, s, dvt, P.invperm.as<int32_t>(),
                           P.nnz, per, dvals);
        TRSPMM_HIP(hipGetLastError());
    }
    return TRSPMM_OK;
}

}  // namespace

int csr_transpose(const int32_t* rowptr, const int32_t* col, const float* val, int64_t rows, int64_t nnz,
                  int64_t ncols, int32_t* t_rowptr, int32_t* t_col, float* t_val, int32_t* perm, hipStream_t s) {
    StreamArena A(s);      // temporaries (declared first: back to the cache after the stream is idle)
    TRSPMM_CHECK(rows >= 0 && nnz >= 0 && ncols >= 0, "csr_transpose: negative size");
    TRSPMM_CHECK(nnz < (int64_t)1 << 31 && ncols < (int64_t)1 << 31 && rows < (int64_t)1 << 31,
                 "csr_transpose: sizes must fit int32");
    TRSPMM_CHECK(t_rowptr && (nnz == 0 || (rowptr && col && t_col && perm)), "csr_transpose: null pointer");
    int rc;
    if (nnz == 0) {
        hipLaunchKernelGGL(k_transpose_rowptr, dim3(nblk(ncols + 1)), dim3(256), 0, s, (const int32_t*)nullptr,
                           (int64_t)0, ncols, t_rowptr);
        TRSPMM_HIP(hipGetLastError());
        return TRSPMM_OK;
    }
    DevBuf err, keys_sorted, ids;
    if ((rc = err.alloc(A, 4)) || (rc = keys_sorted.alloc(A, 4 * nnz)) || (rc = ids.alloc(A, 4 * nnz))) return rc;
    TRSPMM_HIP(hipMemsetAsync(err.p, 0, 4, s));
    hipLaunchKernelGGL(k_check_cols, dim3(nblk(nnz)), dim3(256), 0, s, col, nnz, ncols, err.as<int32_t>());
    TRSPMM_HIP(hipGetLastError());
    int32_t herr = 0;
    TRSPMM_HIP(hipMemcpyAsync(&herr, err.p, 4, hipMemcpyDeviceToHost, s));
    TRSPMM_HIP(hipStreamSynchronize(s));
    TRSPMM_CHECK(herr == 0, "csr_transpose: column index out of range [0, ncols)");
    hipLaunchKernelGGL(k_iota, dim3(nblk(nnz)), dim3(256), 0, s, ids.as<int32_t>(), nnz);
    TRSPMM_HIP(hipGetLastError());
    int bits = 1;
    while (bits < 31 && ((int64_t)1 << bits) < ncols) ++bits;
    size_t tmp_bytes = 0;
    TRSPMM_HIP(rocprim::radix_sort_pairs(nullptr, tmp_bytes, col, keys_sorted.as<int32_t>(), ids.as<int32_t>(), perm,
                                         (size_t)nnz, 0, bits, s));
    DevBuf tmp;
    if ((rc = tmp.alloc(A, tmp_bytes ? tmp_bytes : 1))) return rc;
    TRSPMM_HIP(rocprim::radix_sort_pairs(tmp.p, tmp_bytes, col, keys_sorted.as<int32_t>(), ids.as<int32_t>(), perm,
                                         (size_t)nnz, 0, bits, s));
    hipLaunchKernelGGL(k_transpose_fill, dim3(nblk(nnz)), dim3(256), 0, s, rowptr, rows, val, perm, nnz, t_col, t_val);
    TRSPMM_HIP(hipGetLastError());
    hipLaunchKernelGGL(k_transpose_rowptr, dim3(nblk(ncols + 1)), dim3(256), 0, s, keys_sorted.as<int32_t>(), nnz,
                       ncols, t_rowptr);
    TRSPMM_HIP(hipGetLastError());
    // the temporaries are freed on return: finish the stream's use of them first
    TRSPMM_HIP(hipStreamSynchronize(s));
    return TRSPMM_OK;
}

int launch_sddmm(const int32_t* rowptr, int64_t rows, const int32_t* col, int64_t nnz, int64_t ncols, const void* X,
                 int64_t ldx, const void* Y, int64_t ldy, int64_t K, int dtype, float* out, hipStream_t s) {
    TRSPMM_CHECK(rows >= 0 && nnz >= 0 && ncols >= 0, "sddmm: negative size");
    TRSPMM_CHECK(K > 0 && K % 8 == 0 && ldx % 8 == 0 && ldy % 8 == 0 && ldx >= K && ldy >= K,
                 "sddmm: K and the row strides must be multiples of 8 with stride >= K");
    TRSPMM_CHECK(((uintptr_t)X % 16) == 0 && ((uintptr_t)Y % 16) == 0, "sddmm: X and Y must be 16-byte aligned");
    TRSPMM_CHECK(dtype == TRSPMM_F16 || dtype == TRSPMM_BF16, "sddmm: dtype must be TRSPMM_F16 or TRSPMM_BF16");
    TRSPMM_CHECK(nnz < (int64_t)1 << 31, "sddmm: nnz must fit int32");
    if (nnz == 0) return TRSPMM_OK;
    TRSPMM_CHECK(rowptr && col && out, "sddmm: null pointer");
    if (ncols == 0) {                               // no column is in range: every edge gets 0
        TRSPMM_HIP(hipMemsetAsync(out, 0, 4 * nnz, s));
        return TRSPMM_OK;
    }
    TRSPMM_CHECK(X && Y, "sddmm: null pointer");
    if (dtype == TRSPMM_F16) return launch_sddmm_typed<F16>(rowptr, rows, col, nnz, ncols, X, ldx, Y, ldy, K, out, s);
    return launch_sddmm_typed<BF16>(rowptr, rows, col, nnz, ncols, X, ldx, Y, ldy, K, out, s);
}

static int build_routed(const int32_t* t_rowptr, const int32_t* t_col, const float* t_val, const int32_t* perm,
                        int64_t rows, int64_t nnz, int64_t ncols, int dtype, hipStream_t s, BwdPlan& P);

int bwd_plan_build(const int32_t* t_rowptr, const int32_t* t_col, const float* t_val, const int32_t* perm,
                   int64_t rows, int64_t nnz, int64_t ncols, int dtype, hipStream_t s, BwdPlan& P) {
    StreamArena A(s);      // temporaries (declared first: back to the cache after the stream is idle)
    TRSPMM_CHECK(rows >= 0 && nnz >= 0 && ncols >= 0, "bwd plan: negative size");
    TRSPMM_CHECK(nnz < (int64_t)1 << 31 && rows < (int64_t)1 << 31, "bwd plan: sizes must fit int32");
    TRSPMM_CHECK(dtype == TRSPMM_F16 || dtype == TRSPMM_BF16, "bwd plan: dtype must be TRSPMM_F16 or TRSPMM_BF16");
    TRSPMM_CHECK(t_rowptr && (nnz == 0 || (t_col && t_val && perm)), "bwd plan: null pointer");
    P.rows = rows;
    P.nnz = nnz;
    P.ncols = ncols;
    P.dtype = dtype;
    std::vector<int32_t> rp((size_t)rows + 1);
    TRSPMM_HIP(hipMemcpyAsync(rp.data(), t_rowptr, 4 * (rows + 1), hipMemcpyDeviceToHost, s));
    TRSPMM_HIP(hipStreamSynchronize(s));
    TRSPMM_CHECK(rp[0] == 0 && rp[rows] == nnz, "bwd plan: rowptr must run from 0 to nnz");
    std::vector<int4> items;
    std::vector<int32_t> zero;
    int64_t r = 0;
    while (r < rows) {
        const int32_t len = rp[r + 1] - rp[r];
        TRSPMM_CHECK(len >= 0, "bwd plan: rowptr must be non-decreasing");
        if (len > BWD_ROW_CAP) {                        // long row: chunks with atomic adds
            zero.push_back((int32_t)r);
            for (int32_t b = rp[r]; b < rp[r + 1]; b += BWD_ROW_CAP)
                items.push_back(int4{1, (int32_t)r, b, b + BWD_ROW_CAP < rp[r + 1] ? b + BWD_ROW_CAP : rp[r + 1]});
            ++r;
            continue;
        }
        int64_t r1 = r + 1;                             // a run of whole rows
        while (r1 < rows && r1 - r < BWD_ITEM_ROWS && rp[r1 + 1] - rp[r1] <= BWD_ROW_CAP &&
               rp[r1 + 1] - rp[r] <= BWD_ITEM_CAP)
            ++r1;
        items.push_back(int4{0, (int32_t)r, (int32_t)r1, 0});
        r = r1;
    }
    // Heaviest items first, by the class floor(log2(cost)) with cost = entries
    // + 16 per row (each row loads its B row and stores its dB row), stably so
    // that a class keeps row order.  Blocks go to the 8 XCDs round robin, so the
    // hub rows' chunks, which sit together at low row ids on power-law graphs,
    // start first and spread over all XCDs.
    {
        std::vector<uint8_t> cls(items.size());
        int64_t count[64] = {};
        for (size_t k = 0; k < items.size(); ++k) {
            const int4 x = items[k];
            const uint64_t cost = x.x == 1 ? (uint64_t)(x.w - x.z) + 16
                                           : (uint64_t)(rp[x.z] - rp[x.y]) + 16 * (uint64_t)(x.z - x.y);
            cls[k] = (uint8_t)__builtin_clzll(cost);     // smaller = heavier
            ++count[cls[k]];
        }
        int64_t start[64], run = 0;
        for (int c = 0; c < 64; ++c) {
            start[c] = run;
            run += count[c];
        }
        std::vector<int4> sorted(items.size());
        for (size_t k = 0; k < items.size(); ++k) sorted[start[cls[k]]++] = items[k];
        items.swap(sorted);
    }
    P.n_items = (int64_t)items.size();
    P.n_zero = (int64_t)zero.size();
    int rc;
    const int64_t n1 = nnz > 0 ? nnz : 1;
    if ((rc = P.rowptr.alloc_cached(4 * (rows + 1))) || (rc = P.col.alloc_cached(4 * n1)) || (rc = P.val.alloc_cached(2 * n1)) ||
        (rc = P.invperm.alloc_cached(4 * n1)) || (rc = P.dvals_t.alloc_cached(4 * n1)) ||
        (rc = P.items.alloc_cached(sizeof(int4) * (items.size() ? items.size() : 1))) ||
        (rc = P.zero_rows.alloc_cached(4 * (zero.size() ? zero.size() : 1))))
        return rc;
    TRSPMM_HIP(hipMemcpyAsync(P.rowptr.p, t_rowptr, 4 * (rows + 1), hipMemcpyDeviceToDevice, s));
    if (!items.empty())
        TRSPMM_HIP(hipMemcpyAsync(P.items.p, items.data(), sizeof(int4) * items.size(), hipMemcpyHostToDevice, s));
    if (!zero.empty())
        TRSPMM_HIP(hipMemcpyAsync(P.zero_rows.p, zero.data(), 4 * zero.size(), hipMemcpyHostToDevice, s));
    if (nnz > 0) {
        // columns of A^T index dY's rows: all in [0, ncols); perm must be a
        // permutation of [0, nnz) (entries in range, and the inverse built from
        // it maps every entry back to its own position)
        DevBuf err;
        if ((rc = err.alloc(A, 8))) return rc;
        TRSPMM_HIP(hipMemsetAsync(err.p, 0, 8, s));
        hipLaunchKernelGGL(k_check_cols, dim3(nblk(nnz)), dim3(256), 0, s, t_col, nnz, ncols, err.as<int32_t>());
        TRSPMM_HIP(hipGetLastError());
        hipLaunchKernelGGL(k_check_cols, dim3(nblk(nnz)), dim3(256), 0, s, perm, nnz, nnz, err.as<int32_t>() + 1);
        TRSPMM_HIP(hipGetLastError());
        int32_t bad[2] = {0, 0};
        TRSPMM_HIP(hipMemcpyAsync(bad, err.p, 8, hipMemcpyDeviceToHost, s));
        TRSPMM_HIP(hipStreamSynchronize(s));
        TRSPMM_CHECK(bad[0] == 0, "bwd plan: a column index lies outside [0, ncols)");
        TRSPMM_CHECK(bad[1] == 0, "bwd plan: perm is not a permutation of [0, nnz)");
        TRSPMM_HIP(hipMemcpyAsync(P.col.p, t_col, 4 * nnz, hipMemcpyDeviceToDevice, s));
        hipLaunchKernelGGL(k_to16, dim3(nblk(nnz)), dim3(256), 0, s, t_val, nnz, dtype, P.val.as<uint16_t>());
        TRSPMM_HIP(hipGetLastError());
        hipLaunchKernelGGL(k_invert, dim3(nblk(nnz)), dim3(256), 0, s, perm, nnz, P.invperm.as<int32_t>());
        TRSPMM_HIP(hipGetLastError());
        hipLaunchKernelGGL(k_check_inverse, dim3(nblk(nnz)), dim3(256), 0, s, perm, P.invperm.as<int32_t>(), nnz,
                           err.as<int32_t>());
        TRSPMM_HIP(hipGetLastError());
        TRSPMM_HIP(hipMemcpyAsync(bad, err.p, 4, hipMemcpyDeviceToHost, s));
        TRSPMM_HIP(hipStreamSynchronize(s));
        TRSPMM_CHECK(bad[0] == 0, "bwd plan: perm is not a permutation of [0, nnz)");
    }
    TRSPMM_HIP(hipStreamSynchronize(s));               // host vectors go out of scope
    return build_routed(t_rowptr, t_col, t_val, perm, rows, nnz, ncols, dtype, s, P);
}

// The routed backward: A^T routed into TC sub-blocks and residual records (no
// 2:4 tiles: the fused value gradient covers those two routes), its rows
// reordered when that gathers fewer dY rows, its plan with edge maps, and
// pos[e] = the slot of D (records, then TC values) that holds A's edge e.
// A column of A^T that two rows of a window share goes to a TC sub-block
// (TC density 2), so one gathered dY row serves every row that has it; R-MAT
// scale 18: 6.15 M gathered rows against 8.39 M with every nonzero a record.
// TRSPMM_BACKWARD=lanes in the environment keeps the lane-group kernel instead.
static int build_routed(const int32_t* t_rowptr, const int32_t* t_col, const float* t_val, const int32_t* perm,
                        int64_t rows, int64_t nnz, int64_t ncols, int dtype, hipStream_t s, BwdPlan& P) {
    StreamArena A(s);      // temporaries (declared first: back to the cache after the stream is idle)
    const char* env = std::getenv("TRSPMM_BACKWARD");
    if ((env && std::strcmp(env, "lanes") == 0) || nnz == 0 || rows == 0 || ncols == 0) return TRSPMM_OK;
    trspmm_route_cfg cfg{};
    cfg.part_t = 32;
    cfg.part_c = 32;
    cfg.short_c = 3;
    cfg.density = 2;
    cfg.window = 16;
    cfg.wide = 16;
    cfg.tc_min_util = 0.f;
    cfg.skip_tc = 0;
    cfg.unit_min_nnz = 2;
    cfg.mode = TRSPMM_MODE_TUNED;
    cfg.sptc_min_nnz = 33;
    cfg.flags = TRSPMM_ROUTE_EDGES | TRSPMM_ROUTE_REORDER;
    int rc;
    {
        Meta M;
        if ((rc = route_device(t_rowptr, t_col, t_val, rows, nnz, ncols, cfg, s, M))) return rc;
        P.at.opts.residual_engine = TRSPMM_ENGINE_MFMA;
        // work items of up to 4096 residual entries and 128 TC sub-blocks: on
        // R-MAT scale 18 at K = 256 the fused backward took 437 us against
        // 448-455 us with the forward's defaults (2048, 64); 1024 entries took
        // 477 us, 8192 (hub windows too coarse) 676 us
        P.at.opts.res_cap = 4096;
        P.at.opts.tc_cap = 128;
        const EdgeMaps em = meta_edges(M);
        if ((rc = build_plan(meta_view(M), rows, ncols, dtype, s, P.at, &em, M.row_order, M.rows))) return rc;
    }
    if (!fused_sddmm_supported(P.at.dev) || P.at.dev.n_rval + P.at.dev.n_tval != nnz) return TRSPMM_OK;
    DevBuf cnt;
    if ((rc = P.pos.alloc_cached(4 * nnz)) || (rc = P.perm.alloc_cached(4 * nnz)) || (rc = P.D.alloc_cached(4 * nnz)) ||
        (rc = cnt.alloc(A, 4)))
        return rc;
    TRSPMM_HIP(hipMemsetAsync(P.pos.p, 0xff, 4 * nnz, s));
    TRSPMM_HIP(hipMemsetAsync(cnt.p, 0, 4, s));
    TRSPMM_HIP(hipMemcpyAsync(P.perm.p, perm, 4 * nnz, hipMemcpyDeviceToDevice, s));
    const int32_t* rec_edge = P.at.rec_edge.p ? P.at.rec_edge.as<int32_t>() : P.at.res_edge.as<int32_t>();
    const int64_t n_rec = P.at.dev.n_rval, n_tc = P.at.dev.n_tval;
    if (n_rec > 0) {
        hipLaunchKernelGGL(k_pos_from_records, dim3(nblk(n_rec)), dim3(256), 0, s, rec_edge, n_rec, perm, nnz,
                           (int64_t)0, P.pos.as<int32_t>());
        TRSPMM_HIP(hipGetLastError());
    }
    if (n_tc > 0) {
        hipLaunchKernelGGL(k_pos_from_records, dim3(nblk(n_tc)), dim3(256), 0, s, P.at.tc_edge.as<int32_t>(), n_tc,
                           perm, nnz, n_rec, P.pos.as<int32_t>());
        TRSPMM_HIP(hipGetLastError());
    }
    hipLaunchKernelGGL(k_count_negative, dim3(nblk(nnz)), dim3(256), 0, s, P.pos.as<int32_t>(), nnz, cnt.as<int32_t>());
    TRSPMM_HIP(hipGetLastError());
    int32_t missing = 0;
    TRSPMM_HIP(hipMemcpyAsync(&missing, cnt.p, 4, hipMemcpyDeviceToHost, s));
    TRSPMM_HIP(hipStreamSynchronize(s));
    P.routed = missing == 0;
    return TRSPMM_OK;
}

// new values of A (A's edge order): val[invperm[e]] = vals[e] in the plan's type
int bwd_plan_set_values(BwdPlan& P, const float* vals, int64_t nnz, hipStream_t s) {
    TRSPMM_CHECK(nnz == P.nnz, "bwd plan set_values: nnz must be the plan's nonzero count");
    if (nnz == 0) return TRSPMM_OK;
    TRSPMM_CHECK(vals != nullptr, "bwd plan set_values: vals must not be null");
    hipLaunchKernelGGL(k_scatter16, dim3(nblk(nnz)), dim3(256), 0, s, vals, P.invperm.as<int32_t>(), nnz, P.dtype,
                       P.val.as<uint16_t>());
    TRSPMM_HIP(hipGetLastError());
    if (P.at.has_edges) return plan_set_values(P.at, vals, nnz, P.perm.as<int32_t>(), nnz, s);
    return TRSPMM_OK;
}

int launch_bwd(const BwdPlan& P, const void* dY, int64_t lddy, const void* B, int64_t ldb, int64_t K, float* dB,
               int64_t lddb, float* dvals, hipStream_t s) {
    if (P.routed) {
        // dB = A^T dY on A^T's routed plan, the value gradient of every record on
        // the side, then into A's edge order
        TRSPMM_CHECK(dvals, "fused backward: null pointer");
        int rc = launch_spmm_sddmm(P.at.dev, dY, lddy, K, dB, lddb, B, ldb, P.rows, P.D.as<float>(), s);
        if (rc) return rc;
        const int64_t per = ((P.nnz + kGatherSpan - 1) / kGatherSpan + 7) / 8;
        hipLaunchKernelGGL(k_gather_f32, dim3((unsigned)(8 * per)), dim3(256), 0, s, P.D.as<float>(),
                           P.pos.as<int32_t>(), P.nnz, per, dvals);
        TRSPMM_HIP(hipGetLastError());
        return TRSPMM_OK;
    }
    TRSPMM_CHECK(K > 0 && K % 8 == 0 && K <= 512, "fused backward: K must be a multiple of 8 in [8, 512]");
    TRSPMM_CHECK(lddy % 8 == 0 && ldb % 8 == 0 && lddy >= K && ldb >= K && lddb % 4 == 0 && lddb >= K,
                 "fused backward: row strides must be >= K, dY/B strides multiples of 8, dB stride of 4");
    TRSPMM_CHECK(((uintptr_t)dY % 16) == 0 && ((uintptr_t)B % 16) == 0 && ((uintptr_t)dB % 16) == 0,
                 "fused backward: dY, B and dB must be 16-byte aligned");
    TRSPMM_CHECK(P.rows == 0 || dB, "fused backward: dB must not be null");
    TRSPMM_CHECK(P.nnz == 0 || (dY && B && dvals), "fused backward: null pointer");
    if (P.dtype == TRSPMM_F16) return launch_bwd_typed<F16>(P, dY, lddy, B, ldb, K, dB, lddb, dvals, s);
    return launch_bwd_typed<BF16>(P, dY, lddy, B, ldb, K, dB, lddb, dvals, s);
}

}  // namespace trspmm
