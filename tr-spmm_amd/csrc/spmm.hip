// Routed SpMM for gfx950: C = A_routed · B with FP32 accumulation.
//
// Reference semantics (TR-source/SpMM/mGCNkernel_online.cu):
//   TC route     :54-241  C[16w+r, f] += sum_q A_tc[r, q] * B[t_column[8s+q], f]
//   2:4 route    :812-933 C[16w+r, f] += sum over the tile's 16 columns
//   residual     mGCNkernel.cu:977-1171 with spmm_utils/compute.h:53-87
//   add kernel   :1314-1324
// The reference runs TC, residual-long, residual-short and 2:4 on four streams
// (TC and residual accumulating in fp16, 2:4 into a second buffer) and adds
// the buffers.  Here one multiply is at most three launches on one stream:
//   k_zero_rows   zeroes the rows of the few split (hub) windows,
//   k_tile_routes one wave64 per work item: the dense 16x8 sub-blocks on
//                 v_mfma_f32_16x16x32 and the 2:4 tiles on v_smfmac_f32_16x16x64,
//                 B rows gathered into LDS (next step's rows prefetched into
//                 registers while the current step computes) and read back
//                 transposed with ds_read_b64_tr_b16; writes the window's
//                 16 x K tile,
//   k_residual    (vector-ALU residual engine; the default engine runs the
//                 residual entries inside k_tile_routes as one-hot MFMA
//                 fragments) one wave64 per residual work item: the item's
//                 entries streamed flat over its rows, 16-byte B loads per
//                 lane (8 per lane in flight), fp32 FMA in registers, a row's
//                 sums flushed into the wave's 16-row LDS tile with LDS float
//                 atomics when the row changes; the tile is then stored (or
//                 added to rows the tile kernel wrote).
// Every output element is written by exactly one work item except in split
// windows, whose rows are zeroed and accumulated with fp32 atomics.
#include <algorithm>
#include <atomic>
#include <cstdlib>
#include <type_traits>

#include "plan.h"

namespace trspmm {

namespace {

typedef uint16_t u16x8 __attribute__((ext_vector_type(8)));
typedef short s16x4 __attribute__((__vector_size__(8)));
typedef short s16x8 __attribute__((ext_vector_type(8)));
typedef short s16x16 __attribute__((ext_vector_type(16)));

template <typename DT>
__device__ inline f32x4 mma_dense(typename DT::x8 a, typename DT::x8 b, f32x4 c);
template <>
__device__ inline f32x4 mma_dense<F16>(f16x8 a, f16x8 b, f32x4 c) {
    return __builtin_amdgcn_mfma_f32_16x16x32_f16(a, b, c, 0, 0, 0);
}
template <>
__device__ inline f32x4 mma_dense<BF16>(bf16x8 a, bf16x8 b, f32x4 c) {
    return __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, c, 0, 0, 0);
}

template <typename DT>
__device__ inline f32x4 mma_sparse(typename DT::x8 a, typename DT::x16 b, f32x4 c, int idx);
template <>
__device__ inline f32x4 mma_sparse<F16>(f16x8 a, f16x16 b, f32x4 c, int idx) {
    return __builtin_amdgcn_smfmac_f32_16x16x64_f16(a, b, c, idx, 0, 0);
}
template <>
__device__ inline f32x4 mma_sparse<BF16>(bf16x8 a, bf16x16 b, f32x4 c, int idx) {
    return __builtin_amdgcn_smfmac_f32_16x16x64_bf16(a, b, c, idx, 0, 0);
}

template <typename DT>
__device__ inline float to_f32(uint16_t bits) {
    if constexpr (DT::code == TRSPMM_BF16) return __builtin_bit_cast(float, (uint32_t)bits << 16);
    else return (float)__builtin_bit_cast(_Float16, bits);
}

// LDS transpose read (gfx950 ds_read_b64_tr_b16).  Within each group of 16
// lanes, lane L supplies the address of 4 consecutive 16-bit values (chunk_L)
// and lane j receives element i = chunk_{4i + j/4}[j % 4].  With row-major
// staged B rows, lane L pointing at row (base + L/4), features
// (f + 4*(L%4) .. +4), lane j receives rows base..base+3 of feature f + j:
// four consecutive K entries of one MFMA B-operand column in one instruction.
__device__ inline s16x4 lds_tr16(const uint16_t* p) {
    return __builtin_amdgcn_ds_read_tr16_b64_v4i16((__attribute__((address_space(3))) s16x4*)p);
}

// ============================================================ tile routes
//
// A step is one matrix instruction's worth of K: 4 TC sub-blocks (32 gathered
// rows), 4 2:4 tiles (64 gathered rows) or, with the matrix-core residual
// engine, 32 residual entries.  Lane l gathers the 16-byte pieces
// c = l + 64k of the step's rows (row c / (F/8), piece c % (F/8)).
//
// Software pipeline.  Vector loads return in issue order and a value is waited
// for where it is first used, so every load here is issued well before its
// first use and nothing consumes a loaded value early:
//   * gathered pieces are kept raw with a validity mask applied only when they
//     are parked in LDS, one sub-step after they were issued;
//   * a step's column indices and its A fragment are loaded at the start of
//     the previous step, ahead of that step's gathers.
template <int F>
struct Gather {
    static constexpr int PIECES = F / 8;
    static constexpr int MAXK = PIECES;              // pieces per lane for a 64-row step
    u32x4 v[MAXK];
    uint32_t ok;                                     // bit k: piece k holds data (else parked as zeros)
};

// column indices of a step's staged rows, loaded ahead of its gathers: lane l
// holds the column of row l (R <= 64 rows per step)
template <int F>
struct StepCols {
    int32_t cl;
    uint32_t hi;       // residual steps: the high word of row l's entry record (value | local row << 16)
    int32_t nrows;
};

// cols[row * STRIDE] is the column of staged row `row` (STRIDE 2: the low word
// of a residual entry record); nrows >= 1.  One 4-byte load per lane: the
// gathers' lanes take their rows' columns from it with lane shuffles.  (Each
// lane loading the column of every row it gathers read 8 lanes' worth of the
// same word per row, and the kernel is bound by the vector-memory data it
// gets back.)
template <int F, int R>
__device__ inline void load_cols(StepCols<F>& sc, const int32_t* __restrict__ cols, int nrows, int lane) {
    static_assert(R <= 64, "a step stages at most 64 rows");
    sc.nrows = nrows;
    sc.cl = cols[lane < nrows ? lane : nrows - 1];     // clamped: unconditional load
}

// residual steps: lane l loads entry record l whole (col | value << 32 |
// local row << 48); its column stages the row, and the one-hot A fragment of
// the step is built from the records by lane shuffles
template <int F, int R>
__device__ inline void load_records(StepCols<F>& sc, const uint64_t* __restrict__ rec, int nrows, int lane) {
    static_assert(R <= 64, "a step stages at most 64 rows");
    sc.nrows = nrows;
    const uint64_t x = rec[lane < nrows ? lane : nrows - 1];
    sc.cl = (int32_t)(uint32_t)x;
    sc.hi = (uint32_t)(x >> 32);
}

// B rows addressed with 32-bit element offsets (the launcher checks that B
// holds fewer than 2^31 elements)
template <int F, int R>
__device__ inline void gather_rows(Gather<F>& g, const StepCols<F>& sc, const uint16_t* __restrict__ B, uint32_t ldb,
                                   int f0, int fw, int lane) {
    constexpr int PIECES = F / 8;
    uint32_t ok = 0;
#pragma unroll
    for (int k = 0; k < R * PIECES / 64; ++k) {
        const int c = lane + 64 * k;
        const int row = c / PIECES, piece = c % PIECES;
        const int32_t col = __shfl(sc.cl, row, 64);
        // bitwise & (no short-circuit): a branch per piece would put a wait in each
        const bool v = (row < sc.nrows) & (col >= 0) & (piece * 8 < fw);
        const uint32_t off = v ? (uint32_t)col * ldb + (uint32_t)(f0 + piece * 8) : 0u;
        g.v[k] = *(const u32x4*)(B + off);
        ok |= (uint32_t)v << k;
    }
    g.ok = ok;
}

template <int F, int R, int LDS_STRIDE>
__device__ inline void park_rows(const Gather<F>& g, uint16_t* __restrict__ Bs, int lane) {
    constexpr int PIECES = F / 8;
#pragma unroll
    for (int k = 0; k < R * PIECES / 64; ++k) {
        const int c = lane + 64 * k;
        const int row = c / PIECES, piece = c % PIECES;
        const u32x4 x = ((g.ok >> k) & 1u) ? g.v[k] : u32x4{0u, 0u, 0u, 0u};
        u32x2* dst = (u32x2*)(Bs + row * LDS_STRIDE + piece * 8);
        dst[0] = u32x2{x.x, x.y};
        dst[1] = u32x2{x.z, x.w};
    }
}

// A fragment loaded ahead of its use and applied (masked) there
struct Frag {
    u32x4 a;
    int32_t idx;
    bool ok;
};

// Step loop of one route kind over the steps [0, n) of a work item, for one
// pass of NS slabs (F features each, from feature fp; NS is a template
// argument so that no load sits behind a run-time condition).  KIND 0 = TC
// sub-blocks (32 rows, dense A rows, MFMA), 1 = 2:4 tiles (64 rows, SMFMAC),
// 2 = residual entries as one-hot fragments (32 rows, MFMA).
//
// Sub-step s is (step s / NS, slab s % NS).  The loop is unrolled by two so
// that a sub-step's parity is known at compile time:
//   * even sub-steps accumulate into acc[0] and odd ones into acc[1] (one slab
//     per parity when NS == 2; two partial sums of the same slab when NS == 1,
//     added by the caller);
//   * column indices and A fragments live in a two-entry ring: sub-step s
//     parks its pieces (gathered during s - 1), applies the fragment loaded at
//     s - 2, issues the gathers of s + 1 from the columns loaded at s - 1 and
//     loads the columns and fragment of s + 2.  With NS == 2 the two sub-steps
//     of a step share its columns and fragment: only even sub-steps load them
//     (into entry 0) and the odd one reuses the fragment it applied, which
//     saves the data those loads would return (the kernel is bound by it).
// Every load is unconditional (indices clamped to the last step, fragments of
// sub-steps past the end masked to zero), so the count of loads in flight is
// the same on every path and each wait covers exactly the data it needs.  An
// odd sub-step count gets one masked extra sub-step.
template <int KIND, int F, typename DT, int LDS_STRIDE, int NS>
__device__ inline void run_steps(const PlanDev& p, const uint16_t* __restrict__ B, uint32_t ldb, int K, int fp,
                                 int32_t first, int32_t end, int n, uint16_t* __restrict__ Bs,
                                 const uint16_t* __restrict__ bp, int lane, f32x4 (&acc)[2][F / 16]) {
    using X8 = typename DT::x8;
    using X16 = typename DT::x16;
    constexpr int NB = F / 16;
    constexpr int R = KIND == 1 ? 64 : 32;        // staged rows per step
    constexpr int UNITS = KIND == 2 ? 32 : 4;     // entries / sub-blocks / tiles per step
    const int r = lane & 15, kg = lane >> 4;
    constexpr int sh = NS - 1;                    // step = s >> sh, slab = s & sh
    constexpr bool TWO = NS == 2;                 // sub-steps 2t, 2t + 1 share step t's columns and fragment
    const int nsub = n << sh;

    auto cols_of = [&](int s, StepCols<F>& sc) {
        int step = s >> sh;
        step = step < n ? step : n - 1;
        const int32_t u = first + UNITS * step;
        if constexpr (KIND == 0)
            load_cols<F, R>(sc, p.tc_cols + (int64_t)u * 8, (end - u) * 8 < R ? (end - u) * 8 : R, lane);
        else if constexpr (KIND == 1)
            load_cols<F, R>(sc, p.sp_cols + (int64_t)u * 16, (end - u) * 16 < R ? (end - u) * 16 : R, lane);
        else
            load_records<F, R>(sc, p.res_rec + u, end - u < R ? end - u : R, lane);
    };
    // A fragment of sub-step s (masked at use): TC = row r of sub-block kg of its
    // step; 2:4 = row r of tile kg (4 groups x 2 values) and its index nibbles
    auto frag_of = [&](int s, Frag& fr) {
        int step = s >> sh;
        const bool live = s < nsub;
        step = step < n ? step : n - 1;
        if constexpr (KIND == 0) {
            const int32_t my = first + 4 * step + kg, mc = my < end ? my : end - 1;
            fr.a = *(const u32x4*)(p.tc_dense + ((int64_t)mc * 16 + r) * 4);
            fr.ok = live && my < end;
        } else if constexpr (KIND == 1) {
            const int32_t my = first + 4 * step + kg, mc = my < end ? my : end - 1;
            fr.a = *(const u32x4*)(p.sp_a + ((int64_t)mc * 16 + r) * 4);
            fr.idx = p.sp_idx[(int64_t)mc * 16 + r];
            fr.ok = live && my < end;
        } else {
            fr.ok = live;
        }
    };
    auto gather = [&](Gather<F>& g, const StepCols<F>& sc, int s) {
        const int f0 = fp + (s & sh) * F;
        gather_rows<F, R>(g, sc, B, ldb, f0, K - f0 < F ? K - f0 : F, lane);
    };

    Gather<F> g;
    StepCols<F> cb[2];
    Frag fb[2] = {{{0u, 0u, 0u, 0u}, 0x4444, false}, {{0u, 0u, 0u, 0u}, 0x4444, false}};
    cols_of(0, cb[0]);
    frag_of(0, fb[0]);
    gather(g, cb[0], 0);
    if constexpr (!TWO) {
        cols_of(1, cb[1]);
        frag_of(1, fb[1]);
    }
    X8 a_keep{};
    int idx_keep = 0x4444;

    auto body = [&](int s, auto par_tag) {
        constexpr int PAR = decltype(par_tag)::value;
        constexpr bool FRESH = !TWO || PAR == 0;      // this sub-step starts a step
        X8 a = a_keep;
        int idx = idx_keep;
        if constexpr (KIND == 2 && FRESH) {
            // residual entries as a 16x32 one-hot fragment: staged row k = entry k, and
            // A[r][k] = value of entry k if its local row is r.  Lane (kg, r) takes
            // entries 8kg .. 8kg + 7 of the step from the records that cb[PAR] (this
            // sub-step's, loaded two sub-steps ago) holds one per lane.
            const StepCols<F>& cur = cb[PAR];
            u16x8 abits = {0, 0, 0, 0, 0, 0, 0, 0};
#pragma unroll
            for (int q = 0; q < 8; ++q) {
                const int src = 8 * kg + q;
                const uint32_t hw = (uint32_t)__shfl((int)cur.hi, src, 64);
                const int32_t cq = __shfl(cur.cl, src, 64);
                const bool hit = (s < nsub) & (src < cur.nrows) & ((int)((hw >> 16) & 0xffu) == r) & (cq >= 0);
                abits[q] = hit ? (uint16_t)(hw & 0xffffu) : (uint16_t)0;
            }
            a = __builtin_bit_cast(X8, abits);
            a_keep = a;
        }
        __syncthreads();                               // previous sub-step's LDS reads are done
        park_rows<F, R, LDS_STRIDE>(g, Bs, lane);
        if constexpr (KIND != 2 && FRESH) {
            a = __builtin_bit_cast(X8, fb[PAR].ok ? fb[PAR].a : u32x4{0u, 0u, 0u, 0u});
            idx = fb[PAR].ok ? fb[PAR].idx : 0x4444;   // (0,1) in every group: valid, times zeros
            a_keep = a;
            idx_keep = idx;
        }
        gather(g, TWO ? cb[0] : cb[1 - PAR], s + 1);   // overlaps this sub-step's math
        if constexpr (FRESH) {
            cols_of(s + 2, cb[PAR]);
            frag_of(s + 2, fb[PAR]);
        }
        __syncthreads();
        if constexpr (KIND == 1) {
#pragma unroll
            for (int nb = 0; nb < NB; ++nb) {
                // B operand of v_smfmac_f32_16x16x64 (measured lane layout): slot q of lane
                // (kg, n) holds staged row 32*(q/8) + 8*kg + q%8 of feature n, where staged
                // row 16t + 4g + p is column p of group g of the t-th tile (the A layout:
                // lane (kg, r) = tile kg, row r, groups 0..3, nibble g = p0 | p1 << 2).
                const s16x4 b0 = lds_tr16(bp + nb * 16), b1 = lds_tr16(bp + 4 * LDS_STRIDE + nb * 16);
                const s16x4 b2 = lds_tr16(bp + 32 * LDS_STRIDE + nb * 16);
                const s16x4 b3 = lds_tr16(bp + 36 * LDS_STRIDE + nb * 16);
                const s16x16 bb = {b0[0], b0[1], b0[2], b0[3], b1[0], b1[1], b1[2], b1[3],
                                   b2[0], b2[1], b2[2], b2[3], b3[0], b3[1], b3[2], b3[3]};
                acc[PAR][nb] = mma_sparse<DT>(a, __builtin_bit_cast(X16, bb), acc[PAR][nb], idx);
            }
        } else {
#pragma unroll
            for (int nb = 0; nb < NB; ++nb) {
                // B operand of v_mfma_f32_16x16x32: lane (kg, n) holds staged rows 8kg..8kg+7, feature n
                const s16x4 b0 = lds_tr16(bp + nb * 16), b1 = lds_tr16(bp + 4 * LDS_STRIDE + nb * 16);
                const s16x8 bb = {b0[0], b0[1], b0[2], b0[3], b1[0], b1[1], b1[2], b1[3]};
                acc[PAR][nb] = mma_dense<DT>(a, __builtin_bit_cast(X8, bb), acc[PAR][nb]);
            }
        }
    };
    for (int s = 0; s < nsub; s += 2) {
        body(s, std::integral_constant<int, 0>{});
        body(s + 1, std::integral_constant<int, 1>{});
    }
}

// one work item of the tile kernel: the item's routes over all K features,
// its 16-row tile written at the end (Bs is free again on return).  With a
// ticket (the persistent queue's next index, an atomic's result still in
// flight), the ticket is made uniform before the first pass's stores: every
// load of the pass has been consumed there, so the wait for it costs nothing,
// while waiting after the stores would wait for them too (one in-order
// counter covers loads and stores)
template <int F, typename DT>
__device__ __forceinline__ void tile_item(const PlanDev& p, int64_t it, const uint16_t* __restrict__ B, int64_t ldb,
                                          int K, float* __restrict__ C, int64_t ldc, uint32_t mask, uint16_t* Bs,
                                          int lane, int32_t* ticket) {
    constexpr int LDS_STRIDE = F + 4;     // keeps the 4 lane groups of an operand read on disjoint banks
    constexpr int NB = F / 16;
    const int4 item = p.titems[it];
    const int64_t w = item.x;
    const bool split = item.y != 0;
    int32_t s0 = 0, s1 = 0, t0 = 0, t1 = 0, e0 = 0, e1 = 0;
    if (item.y == 0) {
        s0 = p.tc_sub_off[w], s1 = p.tc_sub_off[w + 1];
        t0 = p.sp_tile_off[w], t1 = p.sp_tile_off[w + 1];
        if (p.res_mfma) e0 = p.res_rowptr[w * 16], e1 = p.res_rowptr[w * 16 + 16];
    } else if (item.y == 1) {
        s0 = item.z, s1 = item.w;
    } else if (item.y == 2) {
        t0 = item.z, t1 = item.w;
    } else {
        e0 = item.z, e1 = item.w;
    }
    if (!(mask & TRSPMM_ROUTE_TC)) s1 = s0;
    if (!(mask & TRSPMM_ROUTE_SPTC)) t1 = t0;
    if (!(mask & TRSPMM_ROUTE_RESIDUAL)) e1 = e0;
    const int r = lane & 15, kg = lane >> 4;
    const int nts = (s1 - s0 + 3) / 4, ntt = (t1 - t0 + 3) / 4, nte = (e1 - e0 + 31) / 32;
    const int64_t row0 = w * 16;
    const uint16_t* bp = Bs + (kg * 8 + (r >> 2)) * LDS_STRIDE + 4 * (r & 3);

    // a pass covers up to two slabs of F features (the step loops run over both,
    // so every step's metadata is read once per pass); acc[0] / acc[1] hold
    // slabs 0 / 1, or two partial sums of slab 0 when the pass has one slab
    for (int fp = 0; fp < K; fp += 2 * F) {
        const int ns = K - fp > F ? 2 : 1;
        f32x4 acc[2][NB];
#pragma unroll
        for (int sl = 0; sl < 2; ++sl)
#pragma unroll
            for (int nb = 0; nb < NB; ++nb) acc[sl][nb] = f32x4{0.f, 0.f, 0.f, 0.f};
        const uint32_t ldb32 = (uint32_t)ldb;
        if (ns == 2) {
            if (nts > 0) run_steps<0, F, DT, LDS_STRIDE, 2>(p, B, ldb32, K, fp, s0, s1, nts, Bs, bp, lane, acc);
            if (ntt > 0) run_steps<1, F, DT, LDS_STRIDE, 2>(p, B, ldb32, K, fp, t0, t1, ntt, Bs, bp, lane, acc);
            if (nte > 0) run_steps<2, F, DT, LDS_STRIDE, 2>(p, B, ldb32, K, fp, e0, e1, nte, Bs, bp, lane, acc);
        } else {
            if (nts > 0) run_steps<0, F, DT, LDS_STRIDE, 1>(p, B, ldb32, K, fp, s0, s1, nts, Bs, bp, lane, acc);
            if (ntt > 0) run_steps<1, F, DT, LDS_STRIDE, 1>(p, B, ldb32, K, fp, t0, t1, ntt, Bs, bp, lane, acc);
            if (nte > 0) run_steps<2, F, DT, LDS_STRIDE, 1>(p, B, ldb32, K, fp, e0, e1, nte, Bs, bp, lane, acc);
        }
        if (ns == 1) {
            // one slab: even and odd sub-steps summed into acc[0]
#pragma unroll
            for (int nb = 0; nb < NB; ++nb) acc[0][nb] += acc[1][nb];
        }
        if (ticket != nullptr && fp == 0) *ticket = __builtin_amdgcn_readfirstlane(*ticket);
        // accumulator element i of lane (kg, n) = C[16w + 4kg + i, f0 + 16nb + n]
        if (!split) {
            // full-line stores: the tile goes through LDS (the staging buffer is free
            // now) and every lane writes 16-byte pieces of whole rows, instead of the
            // 64-byte row fragments of the accumulator layout
            float* T = reinterpret_cast<float*>(Bs);       // [16][2F] fp32 = 8 KB for F = 64
            constexpr int TW = 2 * F;
            static_assert(16 * TW * 4 <= 64 * LDS_STRIDE * 2, "tile must fit the staging buffer");
            __syncthreads();                                // the last sub-step's LDS reads are done
#pragma unroll
            for (int sl = 0; sl < 2; ++sl) {
                if (sl >= ns) continue;
#pragma unroll
                for (int nb = 0; nb < NB; ++nb)
#pragma unroll
                    for (int i = 0; i < 4; ++i) T[(kg * 4 + i) * TW + sl * F + nb * 16 + r] = acc[sl][nb][i];
            }
            __syncthreads();
            const int width = K - fp < ns * F ? K - fp : ns * F;     // features of this pass (multiple of 8)
            const int chunks = width / 4;                            // 16-byte pieces per row
#pragma unroll
            for (int k = 0; k < (16 * TW / 4) / 64; ++k) {
                const int c = lane + 64 * k;
                const int row = c / (TW / 4), piece = c % (TW / 4);
                const int64_t grow = row0 + row;
                if (piece < chunks && grow < p.m_out) {
                    const f32x4 x = *(const f32x4*)(T + row * TW + 4 * piece);
                    __builtin_nontemporal_store(x, (f32x4*)(C + grow * ldc + fp + 4 * piece));
                }
            }
            __syncthreads();                                // T is staging space again for the next pass
        } else {
#pragma unroll
            for (int sl = 0; sl < 2; ++sl) {
                if (sl >= ns) continue;
                const int f0 = fp + sl * F;
                const int fw = K - f0 < F ? K - f0 : F;
#pragma unroll
                for (int nb = 0; nb < NB; ++nb) {
                    const int f = nb * 16 + r;
#pragma unroll
                    for (int i = 0; i < 4; ++i) {
                        const int64_t row = row0 + kg * 4 + i;
                        if (row < p.m_out && f < fw && acc[sl][nb][i] != 0.f)
                            unsafeAtomicAdd(C + row * ldc + f0 + f, acc[sl][nb][i]);
                    }
                }
            }
        }
    }
}

// Work item -> workgroup (PlanDev::item_order):
//   0  one workgroup per item in the plan's item order, heaviest first (the
//      dispatcher deals consecutive workgroups to the 8 XCDs round robin, so
//      the heavy items start first and spread over all XCDs);
//   3  the same with the items in window order;
//   1  one workgroup per item, window order, contiguous item ranges per XCD
//      (xcd_window);
//   2  persistent: a grid sized to the chip's resident capacity (a multiple
//      of 8).  The items (heaviest first) are dealt to 8 queues, queue x
//      holding items x, x + 8, x + 16, ... (the items XCD x runs under
//      order 0); its
//      counter sits on a 256-byte line of its own.  Workgroup b starts on
//      queue b % 8, the queue of the XCD it was dispatched to, takes items
//      with an atomic add (fetching the next index while it works on the
//      current one), and when that queue is empty moves on to the other
//      seven in turn, so an XCD that finishes early takes over the others'
//      tails.  One counter for all items measured slower (about 10 ns per
//      item serialised on one address), and so did tickets of 2 or 4 items
//      (worse balance).  The last workgroup to leave (the done
//      counter) puts every counter back to zero, so the next launch on the
//      stream, or a graph replay, starts from a clean queue without a
//      memset.  Launches of one plan on concurrent streams would share the
//      counters and must not overlap.
//      Order 0 measured faster on every workload timed (R-MAT scale 18,
//      K = 128: 0.241 against 0.287 ms; 1024 batched BA graphs: 0.366 against
//      0.383 ms), so it stays the default and 2 is an autotuner candidate.
//      A likely cause: a wave that ends frees its slot while its stores
//      drain, whereas a persistent wave's next loads share one in-order
//      counter with those stores.
// F = 64 is held to two waves per SIMD (256 VGPRs, no spills); left to
// itself the compiler took a few more registers than that and halved the
// waves.  F = 128 runs one wave per SIMD whatever the bound.
template <int F, typename DT>
__global__ __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(F == 64 ? 2 : 1, F == 64 ? 2 : 1)))
void k_tile_routes(PlanDev p, const uint16_t* __restrict__ B, int64_t ldb, int K, float* __restrict__ C,
                   int64_t ldc, uint32_t mask, int64_t grid) {
    __shared__ __align__(16) uint16_t Bs[64 * (F + 4)];
    const int lane = threadIdx.x;
    if (p.item_order != 2) {
        const int64_t it = p.item_order == 1 ? xcd_window(blockIdx.x, grid) : (int64_t)blockIdx.x;
        if (it < p.n_titems) tile_item<F, DT>(p, it, B, ldb, K, C, ldc, mask, Bs, lane, nullptr);
        return;
    }
    // queue q's next ticket, uniform over the wave; ticket t of queue q is its item q + 8t
    auto take = [&](int q) -> int32_t {
        int32_t t = 0;
        if (lane == 0) t = atomicAdd(p.queue + q * kQueueStride, 1);
        return __builtin_amdgcn_readfirstlane(t);
    };
    int q = blockIdx.x % kXcds;
    int32_t t = take(q);
    for (int visited = 1;;) {
        const int64_t it = q + kXcds * (int64_t)t;
        if (it >= p.n_titems) {                         // queue q is empty: the next one, or leave after all 8
            if (visited == kXcds) break;
            ++visited;
            q = (q + 1) % kXcds;
            t = take(q);
            continue;
        }
        t = 0;
        if (lane == 0) t = atomicAdd(p.queue + q * kQueueStride, 1);      // next ticket, in flight during the item
        tile_item<F, DT>(p, it, B, ldb, K, C, ldc, mask, Bs, lane, &t);
        __syncthreads();                                // the item's last LDS reads are done before the next's writes
    }
    if (lane == 0) {
        __threadfence();
        int32_t* done = p.queue + kXcds * kQueueStride;
        if (atomicAdd(done, 1) == (int32_t)gridDim.x - 1) {
            for (int x = 0; x < kXcds; ++x) atomicExch(p.queue + x * kQueueStride, 0);
            atomicExch(done, 0);
        }
    }
}

// ============================================================ deep-pipelined tile routes
//
// k_routes: the same work items, routes and matrix instructions as
// k_tile_routes, with the gathers reorganised around what bounds the kernel,
// the B-row data a wave has in flight.  Measured with a gather-only kernel
// over R-MAT scale 18's residual entries (K = 128, one wave per window, 8
// waves per CU): 4 KB in flight per wave (k_tile_routes' figure) 0.212 ms,
// 8 KB 0.129 ms, 16 KB 0.121 ms; direct-to-LDS loads of the same rows
// (global_load_lds_dwordx4) were slower than loads into registers at equal
// bytes in flight (0.138 against 0.118 ms), so the rows go through registers:
//   * a pass covers PW = 128 features (64 when K <= 64): a step stages full
//     B rows, 32 (TC sub-blocks, residual entries) or 64 (2:4 tiles) of them;
//   * the next step's gathers (8 KB, or 16 KB for 64-row steps) are in flight
//     while a step is applied (two 4 KB steps at PW = 64);
//   * B is read as a raw buffer: one offset register per load, and masked
//     pieces use an out-of-range offset, which returns zeros without a memory
//     access;
//   * a step's column indices (residual: records) are loaded 2 DEP steps
//     ahead and staged in LDS DEP steps ahead; its A fragment is loaded DEP
//     steps ahead, right after the previous use of its ring slot.  Loads
//     return in issue order, so waiting for them never waits for gathers that
//     are still wanted in flight;
//   * staged rows sit RS = 2 PW + 32 bytes apart and logical row 32 h + 8 kg + j
//     (j < 8) is stored as physical row 32 h + 4 j + kg.  The transposed reads
//     of the MFMA B operand (rows 8 kg + j, j < 4, of all four kg at once) then
//     fall on distinct 32-byte bank windows, the 16-byte parking writes cover
//     each bank once per row, and every address is a per-lane base plus an
//     immediate.
template <int PW>
struct DeepCfg {
    static constexpr int PIECES = PW / 8;                  // 16-byte pieces per staged row
    static constexpr int RS = 2 * PW + 32;                 // LDS bytes per staged row
    // steps in flight for steps of R staged rows: one step of 8 KB (32 rows x 128
    // features; 64-row steps at PW = 128 keep one step of 16 KB), or two steps of
    // 4 KB at PW = 64.  (Gather-only kernel, R-MAT scale 18, 8 waves per CU: 16 KB
    // per wave 0.121 ms, 8 KB 0.129 ms; 32 KB per wave at 4 waves per CU
    // 0.154 ms; the registers of 16 KB in flight do not fit two waves per SIMD
    // beside the step's other state.)
    static constexpr int dep(int R) { return PW >= 128 ? 1 : (R == 32 ? 2 : 1); }
};

__device__ inline int deep_phys(int L) { return (L & ~31) | ((L & 7) << 2) | ((L >> 3) & 3); }

// A step's metadata travels in two hops.  Its column indices (residual: its
// 8-byte records) are loaded into registers 2 DEP steps ahead, one per lane
// (lane l holds staged row l), and written to a small LDS ring DEP steps
// ahead; from there each lane reads the columns of the rows it gathers, and
// (residual) the high words of the 8 entries its one-hot fragment covers,
// all as one base address plus immediates.  Invalid entries reach the ring
// as column -1 and high word 0 (the plan also stores a zero value for a
// record whose column is invalid).
struct DeepRec {
    int32_t cl;
    uint32_t hi;
    float part;        // fused SDDMM, passes after the first: the record's partial sum so far
};
// a step's A fragment (TC, 2:4), loaded right after the fragment of the step
// DEP earlier was applied (ring of DEP); by the time it is needed, the gathers
// issued before it have been waited for, and those issued after it stay in flight
struct DeepFrag {
    u32x4 a;
    int32_t idx;
    bool ok;
};

// Fused SDDMM (residual route of a plan of A^T, every nonzero a record): for
// record k of a step, (lr, slot k) of the window's 16 x 32 block
//     S = X_w G^T,   X_w = the 16 rows of X (rows of A^T, i.e. B of the forward)
// in this pass's features, G = the step's 32 staged rows (rows of dY), is the
// value gradient <X[row], dY[col]> of that nonzero; D[record] receives it (the
// first pass stores, later passes add what the earlier ones stored).  A pass
// covers H PW-feature halves, each step gathered once per half, so K <= 2 PW
// is one pass (H = 2 holds both halves' accumulators).  xa holds the pass's A
// operand (lane (kg, r): X row r, features 32 ks + 8 kg .. + 7).
template <int PW, int H>
struct Sdd {
    u32x4 xa[H * PW / 32];
    float* D;
    bool rmw;
};
struct NoSdd {};

// slot of block h (16 of the step's 32 staged rows) read by lane (kg, r): the
// physical rows of one block's 16 slots are two per residue mod 8, so each
// 16-lane group's 16-byte reads cover the LDS banks twice (2-way, the least
// the 2 PW + 32-byte row stride allows)
__device__ inline int sdd_slot(int h, int r) { return ((r >> 2) << 3) | ((2 * h + ((r >> 1) & 1)) << 1) | (r & 1); }

// staged rows of step s (s clamped to the last step)
template <int KIND>
__device__ inline int deep_rows(int32_t first, int32_t end, int n, int s) {
    const int sc = s < n ? s : n - 1;
    if constexpr (KIND == 2) {
        const int32_t u = first + 32 * sc;
        return end - u < 32 ? end - u : 32;
    } else {
        constexpr int PER = KIND == 0 ? 8 : 16;
        const int32_t u = first + 4 * sc;
        return (end - u) * PER < 4 * PER ? (end - u) * PER : 4 * PER;
    }
}

// 16-byte (or smaller) load at a uniform base plus a 32-bit byte offset (the
// scalar-base addressing form: one offset register per load, no 64-bit math)
template <typename T>
__device__ inline T load_at(const void* base, uint32_t byte_off) {
    return *(const T*)((const uint8_t*)base + byte_off);
}

template <int KIND>
__device__ inline void deep_rec(const PlanDev& p, int32_t first, int32_t end, int n, int s, int lane, DeepRec& m,
                                const float* part = nullptr) {
    const int sc = s < n ? s : n - 1;                      // clamped: the loads stay in range
    const int nrows = deep_rows<KIND>(first, end, n, s);
    const uint32_t l = (uint32_t)(lane < nrows ? lane : nrows - 1);
    if (part) {                                            // lane 32 + j: the partial of row j
        const uint32_t j = (uint32_t)((lane & 31) < nrows ? (lane & 31) : nrows - 1);
        m.part = load_at<float>(part, 4u * ((uint32_t)(first + 32 * sc) + j));
    }
    if constexpr (KIND == 0) {
        m.cl = load_at<int32_t>(p.tc_cols, 4u * ((uint32_t)(first + 4 * sc) * 8u + l));
    } else if constexpr (KIND == 1) {
        m.cl = load_at<int32_t>(p.sp_cols, 4u * ((uint32_t)(first + 4 * sc) * 16u + l));
    } else {
        const uint64_t x = load_at<uint64_t>(p.res_rec, 8u * ((uint32_t)(first + 32 * sc) + l));
        m.cl = (int32_t)(uint32_t)x;
        m.hi = (uint32_t)(x >> 32);
    }
}

// the records of step s into LDS ring slot ms: 8 bytes per lane, lane l's at
// byte 8 l (lanes past the step's rows store column -1, value 0; no branch)
template <int KIND>
__device__ inline void deep_stage(const DeepRec& m, int32_t first, int32_t end, int n, int s, uint8_t* __restrict__ ms,
                                  int lane, bool part = false) {
    const bool v = (s < n) & (lane < deep_rows<KIND>(first, end, n, s));
    // (residual steps have at most 32 rows: bytes 256.. of the slot are free for
    // the partials of the fused SDDMM, row j's at 256 + 8 j)
    if constexpr (KIND == 2)
        *(u32x2*)(ms + 8 * lane) = (part & (lane >= 32)) ? u32x2{__builtin_bit_cast(uint32_t, m.part), 0u}
                                                         : u32x2{(uint32_t)(v ? m.cl : -1), v ? m.hi : 0u};
    else
        *(int32_t*)(ms + 8 * lane) = v ? m.cl : -1;
}

template <int KIND>
__device__ inline void deep_frag(const PlanDev& p, int32_t first, int32_t end, int n, int s, int lane, DeepFrag& f) {
    const int sc = s < n ? s : n - 1;
    const int r = lane & 15, kg = lane >> 4;
    const int32_t my = first + 4 * sc + kg, mc = my < end ? my : end - 1;
    f.ok = (s < n) & (my < end);
    const uint32_t row = (uint32_t)mc * 16u + (uint32_t)r;
    if constexpr (KIND == 0) {
        f.a = load_at<u32x4>(p.tc_dense, 16u * row);
    } else if constexpr (KIND == 1) {
        f.a = load_at<u32x4>(p.sp_a, 16u * row);
        f.idx = load_at<uint16_t>(p.sp_idx, 2u * row);
    }
}

// B as a raw buffer: 16-byte loads at a 32-bit byte offset from one scalar
// descriptor.  An offset at or beyond the buffer's size returns zeros without
// a memory access, which is how masked pieces are staged as zeros (the
// launcher keeps B under 2^31 bytes for this kernel, so 2^31 is always out).
constexpr uint32_t kBufferWord3 = 0x00020000u;      // gfx9-family raw buffer (32-bit data format)
constexpr uint32_t kOutOfRange = 0x80000000u;

// the step's gathers: load k of lane l is piece l % PIECES of staged row
// RPL k + l / PIECES (RPL = 64 / PIECES rows per load), its column read from
// the step's LDS ring slot
template <int PW, int R>
__device__ inline void deep_gather(u32x4 (&g)[R * DeepCfg<PW>::PIECES / 64], const uint8_t* __restrict__ ms,
                                   __amdgpu_buffer_rsrc_t rsB, uint32_t ldb2, int f0, int fw, int lane) {
    constexpr int PIECES = DeepCfg<PW>::PIECES;
    constexpr int RPL = 64 / PIECES;
    const int piece = lane % PIECES;
    const int32_t* mcol = (const int32_t*)(ms + 8 * (lane / PIECES));
    const bool fok = piece * 8 < fw;
    const uint32_t fo = 2u * (uint32_t)(f0 + piece * 8);
#pragma unroll
    for (int k = 0; k < R * PIECES / 64; ++k) {
        const int32_t col = mcol[2 * RPL * k];
        const uint32_t off = (fok & (col >= 0)) ? (uint32_t)col * ldb2 + fo : kOutOfRange;
        g[k] = __builtin_bit_cast(u32x4, __builtin_amdgcn_raw_buffer_load_b128(rsB, off, 0, 0));
    }
    // keeps the loads here: the step loop's early exits would otherwise let the
    // compiler sink them to their first use, a step or two later, which leaves
    // nothing in flight (a compiler barrier only; no wait is emitted for it)
    asm volatile("" ::: "memory");
}

// staged row L (load k, lane l) goes to physical row deep_phys(L): for load k
// that is the lane's row of load 0 plus a constant
template <int PW, int R>
__device__ inline void deep_park(const u32x4 (&g)[R * DeepCfg<PW>::PIECES / 64], uint8_t* __restrict__ Bs, int lane) {
    constexpr int PIECES = DeepCfg<PW>::PIECES;
    constexpr int RPL = 64 / PIECES;
    constexpr int RS = DeepCfg<PW>::RS;
    const int l0 = lane / PIECES;                          // staged row of load 0
    uint8_t* base = Bs + deep_phys(l0) * RS + 16 * (lane % PIECES);
#pragma unroll
    for (int k = 0; k < R * PIECES / 64; ++k) {
        // deep_phys(RPL k + l0) - deep_phys(l0) = deep_phys(RPL k) for RPL = 2, 4 or 8, l0 < RPL
        const int L = RPL * k;
        const int dphys = (L & ~31) + ((L & 7) << 2) + ((L >> 3) & 3);
        *(u32x4*)(base + dphys * RS) = g[k];
    }
}

// Steps [0, n) of one route kind for one pass of PW features from feature
// fp.  Step s: park s (its gathers were issued DEP steps ago), stage the
// records of s + DEP in LDS, load the records of s + 2 DEP, issue the gathers
// of s + DEP into the registers just parked, apply s, then load the A
// fragment of s + DEP into the slot s's fragment used.  The loop is unrolled
// so that every ring index is a compile-time constant.
template <int KIND, int PW, typename DT, int H = 1, typename SD = NoSdd>
__device__ inline void run_deep(const PlanDev& p, __amdgpu_buffer_rsrc_t rsB, uint32_t ldb2, int K, int fp,
                                int32_t first, int32_t end, int n, uint8_t* __restrict__ Bs, uint8_t* __restrict__ Ms,
                                int lane, f32x4 (&acc)[H][PW / 16], const SD& sd = SD{}) {
    constexpr bool FUSED = !std::is_same<SD, NoSdd>::value;
    static_assert(!FUSED || KIND == 0 || KIND == 2, "the fused SDDMM covers TC sub-blocks and residual records");
    static_assert(H == 1 || (FUSED && H == 2), "two halves per step: the fused SDDMM only");
    using X8 = typename DT::x8;
    using X16 = typename DT::x16;
    constexpr int R = KIND == 1 ? 64 : 32;
    constexpr int NL = R * DeepCfg<PW>::PIECES / 64;       // loads per lane per step
    constexpr int DEP = DeepCfg<PW>::dep(R);               // steps in flight
    static_assert(DEP == 1 || DEP == 2, "1 or 2 steps in flight");
    constexpr int MR = 2 * DEP;                            // record / fragment rings (>= DEP + 1)
    constexpr int U = MR;                                  // unroll: a multiple of DEP, MR and H
    constexpr int NB = PW / 16;
    constexpr int RS = DeepCfg<PW>::RS;
    constexpr int MS = 512;                                // bytes per LDS record slot (64 lanes x 8)
    static_assert(U % H == 0, "the unrolled bodies cover whole steps");
    const int r = lane & 15, kg = lane >> 4;
    // virtual step v = H s + half: step s's records, features fp + PW half ..
    auto fw_of = [&](int half) { return K - (fp + PW * half) < PW ? K - (fp + PW * half) : PW; };
    const int nv = H * n;
    // transposed reads: lane (kg, r) of n-block nb reads logical rows 8 kg + r / 4
    // (+ 4, + 32, + 36), features 16 nb + 4 (r % 4) .. + 3
    const uint8_t* rd = Bs + ((r & 12) + kg) * RS + 8 * (r & 3);
    // fused SDDMM: the B operand of block h, slot sdd_slot(h, r), features 32 ks + 8 kg ..
    const uint8_t* sdr0 = Bs + deep_phys(sdd_slot(0, r)) * RS + 16 * kg;
    const uint8_t* sdr1 = Bs + deep_phys(sdd_slot(1, r)) * RS + 16 * kg;
    const float* part = nullptr;                           // residual records: the partials of earlier passes
    if constexpr (FUSED && KIND == 2) part = sd.rmw ? sd.D : nullptr;

    DeepRec m[MR];
    DeepFrag fr[DEP];
    u32x4 g[DEP][NL];
    // prologue: records of steps 0 .. 2 DEP - 1, the first DEP staged, then the
    // first DEP steps' fragments and gathers
#pragma unroll
    for (int j = 0; j < MR; ++j) deep_rec<KIND>(p, first, end, n, j / H, lane, m[j], part);
#pragma unroll
    for (int j = 0; j < DEP; ++j) deep_stage<KIND>(m[j], first, end, n, j / H, Ms + MS * j, lane, part != nullptr);
    if constexpr (KIND != 2) {
#pragma unroll
        for (int j = 0; j < DEP; ++j) deep_frag<KIND>(p, first, end, n, j, lane, fr[j]);
    }
#pragma unroll
    for (int j = 0; j < DEP; ++j)
        deep_gather<PW, R>(g[j], Ms + MS * j, rsB, ldb2, fp + PW * (j % H), fw_of(j % H), lane);
    f32x4 sacc[2];                                         // fused SDDMM: S of the step's two blocks

    auto body = [&](int v, auto par_tag) {
        constexpr int PAR = decltype(par_tag)::value;
        constexpr int GB = PAR % DEP;
        constexpr int NXT = (PAR + DEP) % MR;              // ring slot of virtual step v + DEP
        constexpr int HALF = PAR % H;                      // (U is a multiple of H: v % H == PAR % H)
        constexpr int NHALF = (PAR + DEP) % H;             // half of virtual step v + DEP
        const int s = v / H;                               // the records' step
        __builtin_amdgcn_sched_barrier(0);
        // fused SDDMM over TC sub-blocks: lane (kg, r) writes the value gradients of
        // sub-block r / 4 of the step (its columns sdd_slot(h, r) % 8, rows 4 kg ..
        // 4 kg + 3), which needs the sub-block's occupancy bits and value offset.
        // Loaded here, before this step's record loads and next gathers, so that
        // waiting for them does not wait for loads that are meant to stay in flight.
        [[maybe_unused]] uint64_t tcb_lo = 0, tcb_hi = 0;
        [[maybe_unused]] int32_t tcb_off = 0;
        [[maybe_unused]] const int32_t tcb_sb = first + 4 * s + (r >> 2);
        if constexpr (FUSED && KIND == 0 && HALF == H - 1) {
            const int32_t sbc = tcb_sb < end ? tcb_sb : end - 1;
            tcb_lo = p.tc_bits[2 * sbc];
            tcb_hi = p.tc_bits[2 * sbc + 1];
            tcb_off = p.tc_valoff[sbc];
            asm volatile("" ::: "memory");
        }
        deep_park<PW, R>(g[GB], Bs, lane);
        deep_stage<KIND>(m[NXT], first, end, n, (v + DEP) / H, Ms + MS * NXT, lane, part != nullptr);
        deep_rec<KIND>(p, first, end, n, (v + 2 * DEP) / H, lane, m[(PAR + 2 * DEP) % MR], part);
        deep_gather<PW, R>(g[GB], Ms + MS * NXT, rsB, ldb2, fp + PW * NHALF, fw_of(NHALF), lane);
        X8 a;
        int idx = 0x4444;
        if constexpr (KIND == 2) {
            // one-hot fragment: lane (kg, r) takes entries 8 kg .. 8 kg + 7 of the step
            const uint32_t* hw = (const uint32_t*)(Ms + MS * (PAR % MR) + 64 * kg + 4);
            u16x8 abits = {0, 0, 0, 0, 0, 0, 0, 0};
#pragma unroll
            for (int q = 0; q < 8; ++q) {
                const uint32_t h = hw[2 * q];
                abits[q] = ((h >> 16) & 0xffu) == (uint32_t)r ? (uint16_t)(h & 0xffffu) : (uint16_t)0;
            }
            a = __builtin_bit_cast(X8, abits);
        } else {
            const DeepFrag& cf = fr[GB];
            a = __builtin_bit_cast(X8, cf.ok ? cf.a : u32x4{0u, 0u, 0u, 0u});
            if constexpr (KIND == 1) idx = cf.ok ? cf.idx : 0x4444;     // (0,1) in every group: valid, times zeros
        }
#pragma unroll
        for (int nb = 0; nb < NB; ++nb) {
            const uint16_t* ap = (const uint16_t*)(rd + 32 * nb);
            const s16x4 b0 = lds_tr16(ap), b1 = lds_tr16(ap + 8 * RS);          // logical rows +4: physical +16
            if constexpr (KIND == 1) {
                const s16x4 b2 = lds_tr16(ap + 16 * RS), b3 = lds_tr16(ap + 24 * RS);   // logical +32, +36
                const s16x16 bb = {b0[0], b0[1], b0[2], b0[3], b1[0], b1[1], b1[2], b1[3],
                                   b2[0], b2[1], b2[2], b2[3], b3[0], b3[1], b3[2], b3[3]};
                acc[HALF][nb] = mma_sparse<DT>(a, __builtin_bit_cast(X16, bb), acc[HALF][nb], idx);
            } else {
                const s16x8 bb = {b0[0], b0[1], b0[2], b0[3], b1[0], b1[1], b1[2], b1[3]};
                acc[HALF][nb] = mma_dense<DT>(a, __builtin_bit_cast(X8, bb), acc[HALF][nb]);
            }
        }
        // keep the transposed reads next to the matrix instruction that takes them
        // (hoisted all at once they hold 32 more registers than the step can spare)
#pragma unroll
        for (int nb = 0; nb < NB; ++nb) {
            __builtin_amdgcn_sched_group_barrier(0x100, KIND == 1 ? 4 : 2, 0);   // DS reads
            __builtin_amdgcn_sched_group_barrier(0x008, 1, 0);                    // MFMA
        }
        if constexpr (FUSED) {
            // S = X_w G^T for the step's two 16-slot blocks; lane (kg, r) ends with
            // S[4 kg + i][slot sdd_slot(h, r)], i < 4, and keeps the one of its record's row
            const uint8_t* slot = Ms + MS * (PAR % MR);
#pragma unroll
            for (int h = 0; h < 2; ++h) {
                const uint8_t* sdr = h == 0 ? sdr0 : sdr1;
                if constexpr (HALF == 0) sacc[h] = f32x4{0.f, 0.f, 0.f, 0.f};
#pragma unroll
                for (int ks = 0; ks < PW / 32; ++ks)
                    sacc[h] = mma_dense<DT>(__builtin_bit_cast(X8, sd.xa[HALF * (PW / 32) + ks]),
                                            *(const X8*)(sdr + 64 * ks), sacc[h]);
            }
            if constexpr (HALF == H - 1 && KIND == 0) {
                // S[row][slot] of a set bit is the gradient of the sub-block's value
                // of that rank (values in bit order: rows 0-7 in the low word first)
                if (tcb_sb < end) {
                    const uint64_t bits = kg < 2 ? tcb_lo : tcb_hi;
                    const int32_t base = tcb_off + (kg < 2 ? 0 : __popcll(tcb_lo));
                    float* Dt = sd.D + p.n_rval;
#pragma unroll
                    for (int h = 0; h < 2; ++h) {
                        const int q = sdd_slot(h, r) & 7;
#pragma unroll
                        for (int i = 0; i < 4; ++i) {
                            const int b = ((4 * kg + i) & 7) * 8 + q;
                            if ((bits >> b) & 1ull) {
                                const int32_t idx = base + __popcll(bits & ((1ull << b) - 1ull));
                                float v = sacc[h][i];
                                if (sd.rmw) v += Dt[idx];
                                Dt[idx] = v;
                            }
                        }
                    }
                }
            }
            if constexpr (HALF == H - 1 && KIND == 2) {
#pragma unroll
            for (int h = 0; h < 2; ++h) {
                const f32x4 sa = sacc[h];
                const int L = sdd_slot(h, r);
                const uint32_t hw = *(const uint32_t*)(slot + 8 * L + 4);
                const int i = (int)((hw >> 16) & 0xffu) - 4 * kg;
                float v = i == 0 ? sa[0] : (i == 1 ? sa[1] : (i == 2 ? sa[2] : sa[3]));
                if (part) v += *(const float*)(slot + 256 + 8 * L);
                // into the record's column word (no longer read): the step's 32
                // values then leave in slot order, one 128-byte store
                if ((unsigned)i < 4u) *(float*)(slot + 8 * L) = v;
            }
            const int32_t rec = first + 32 * s + lane;
            const float v = *(const float*)(slot + 8 * (lane & 31));
            if ((lane < 32) & (rec < end)) sd.D[rec] = v;
            }
        }
        // the next step's fragment, once every half of this step has applied it
        if constexpr (KIND != 2 && (PAR % H) == H - 1) deep_frag<KIND>(p, first, end, n, s + DEP, lane, fr[GB]);
    };
    for (int v = 0; v < nv; v += U) {
        body(v, std::integral_constant<int, 0>{});
        if (v + 1 >= nv) break;
        body(v + 1, std::integral_constant<int, 1>{});
        if constexpr (U >= 4) {
            if (v + 2 >= nv) break;
            body(v + 2, std::integral_constant<int, 2>{});
            if (v + 3 >= nv) break;
            body(v + 3, std::integral_constant<int, 3>{});
        }
        if constexpr (U == 8) {
            if (v + 4 >= nv) break;
            body(v + 4, std::integral_constant<int, 4>{});
            if (v + 5 >= nv) break;
            body(v + 5, std::integral_constant<int, 5>{});
            if (v + 6 >= nv) break;
            body(v + 6, std::integral_constant<int, 6>{});
            if (v + 7 >= nv) break;
            body(v + 7, std::integral_constant<int, 7>{});
        }
    }
}

// one work item: passes of PW features; per pass the routes' steps, then the
// 16-row tile written (full lines through LDS, or atomically for split windows)
// fused SDDMM operand of a launch (X = null: none)
struct SddArgs {
    const uint16_t* X;     // [x_rows, ldx] 16-bit, rows indexed like the output rows
    int64_t ldx, x_rows;
    float* D;              // one float per residual record
};

// FH: 0 = the plain multiply; 1 or 2 = with the fused SDDMM, passes of FH halves;
// RM: the plan has a row map (a reordered routing): tile row r goes to output
// row p.row_map[16 w + r]
template <int PW, typename DT, int FH = 0, bool RM = false>
__device__ __forceinline__ void deep_item(const PlanDev& p, int64_t it, const uint16_t* B, int64_t ldb,
                                          int K, float* __restrict__ C, int64_t ldc, uint32_t mask, uint8_t* Bs,
                                          uint8_t* Ms, int lane, const SddArgs& sa = SddArgs{}) {
    constexpr int NB = PW / 16;
    constexpr bool FUSED = FH > 0;
    constexpr int H = FUSED ? FH : 1;
    const int4 item = p.titems[it];
    const int64_t w = item.x;
    const bool split = item.y != 0;
    int32_t s0 = 0, s1 = 0, t0 = 0, t1 = 0, e0 = 0, e1 = 0;
    if (item.y == 0) {
        s0 = p.tc_sub_off[w], s1 = p.tc_sub_off[w + 1];
        t0 = p.sp_tile_off[w], t1 = p.sp_tile_off[w + 1];
        if (p.res_mfma) e0 = p.res_rowptr[w * 16], e1 = p.res_rowptr[w * 16 + 16];
    } else if (item.y == 1) {
        s0 = item.z, s1 = item.w;
    } else if (item.y == 2) {
        t0 = item.z, t1 = item.w;
    } else {
        e0 = item.z, e1 = item.w;
    }
    if (!(mask & TRSPMM_ROUTE_TC)) s1 = s0;
    if (!(mask & TRSPMM_ROUTE_SPTC)) t1 = t0;
    if (!(mask & TRSPMM_ROUTE_RESIDUAL)) e1 = e0;
    const int r = lane & 15, kg = lane >> 4;
    const int nts = (s1 - s0 + 3) / 4, ntt = (t1 - t0 + 3) / 4, nte = (e1 - e0 + 31) / 32;
    const int64_t row0 = w * 16;
    const int rows_out = p.m_out - row0 < 16 ? (int)(p.m_out - row0) : 16;
    float* Crow = C + row0 * ldc;                              // the tile's first row (uniform)
    const __amdgpu_buffer_rsrc_t rsB =
        __builtin_amdgcn_make_buffer_rsrc((void*)B, (short)0, (int)(uint32_t)(2 * p.ncols * ldb), (int)kBufferWord3);
    const uint32_t ldb2 = 2u * (uint32_t)ldb;
    int32_t rmap = -1;                                         // RM: lane l < 16 holds tile row l's output row
    if constexpr (RM) rmap = lane < 16 ? p.row_map[row0 + lane] : -1;
    for (int fp = 0; fp < K; fp += H * PW) {
        f32x4 acc[H][NB];
#pragma unroll
        for (int hf = 0; hf < H; ++hf)
#pragma unroll
            for (int nb = 0; nb < NB; ++nb) acc[hf][nb] = f32x4{0.f, 0.f, 0.f, 0.f};
        // the item's bounds, re-read in each pass: keeps the compiler from hoisting the
        // steps' address arithmetic out of the pass loop into long-lived registers
        int32_t bs0 = s0, bs1 = s1, bt0 = t0, bt1 = t1, be0 = e0, be1 = e1;
        asm volatile("" : "+s"(bs0), "+s"(bs1), "+s"(bt0), "+s"(bt1), "+s"(be0), "+s"(be1));
        if constexpr (FUSED) {
            // the plan holds TC sub-blocks and residual records only (checked by the launcher)
            Sdd<PW, H> sd;
            sd.D = sa.D;
            sd.rmw = fp > 0;
            const __amdgpu_buffer_rsrc_t rsX = __builtin_amdgcn_make_buffer_rsrc(
                (void*)sa.X, (short)0, (int)(uint32_t)(2 * sa.x_rows * sa.ldx), (int)kBufferWord3);
            // X row of tile row r: the plan row, or its output row under a row map
            int64_t xrow = row0 + r;
            if constexpr (RM) xrow = __shfl(rmap, r);
            const bool rok = (xrow >= 0) & (xrow < sa.x_rows);
#pragma unroll
            for (int ks = 0; ks < H * PW / 32; ++ks) {
                const int f = fp + 32 * ks + 8 * kg;
                const uint32_t off = (rok & (f < K)) ? (uint32_t)xrow * (uint32_t)(2 * sa.ldx) + 2u * (uint32_t)f
                                                     : kOutOfRange;
                sd.xa[ks] = __builtin_bit_cast(u32x4, __builtin_amdgcn_raw_buffer_load_b128(rsX, off, 0, 0));
            }
            if (nts > 0) run_deep<0, PW, DT, H>(p, rsB, ldb2, K, fp, bs0, bs1, nts, Bs, Ms, lane, acc, sd);
            if (nte > 0) run_deep<2, PW, DT, H>(p, rsB, ldb2, K, fp, be0, be1, nte, Bs, Ms, lane, acc, sd);
        } else {
            if (nts > 0) run_deep<0, PW, DT>(p, rsB, ldb2, K, fp, bs0, bs1, nts, Bs, Ms, lane, acc);
            if (ntt > 0) run_deep<1, PW, DT>(p, rsB, ldb2, K, fp, bt0, bt1, ntt, Bs, Ms, lane, acc);
            if (nte > 0) run_deep<2, PW, DT>(p, rsB, ldb2, K, fp, be0, be1, nte, Bs, Ms, lane, acc);
        }
#pragma unroll
        for (int hf = 0; hf < H; ++hf) {
        const int fh = fp + PW * hf;                           // this half's first feature
        if (fh >= K) break;
        const int width = K - fh < PW ? K - fh : PW;          // features of this half (multiple of 8)
        // the output stride, re-read here for the same reason as the bounds above
        uint32_t ldc32 = (uint32_t)ldc;
        asm volatile("" : "+s"(ldc32));
        // accumulator element i of lane (kg, n) = C[16w + 4kg + i, fh + 16nb + n]
        if (!split) {
            float* T = reinterpret_cast<float*>(Bs);           // [16][PW] fp32, the staging space is free
#pragma unroll
            for (int nb = 0; nb < NB; ++nb)
#pragma unroll
                for (int i = 0; i < 4; ++i) T[(kg * 4 + i) * PW + nb * 16 + r] = acc[hf][nb][i];
            // store k of lane l writes 16 bytes of row prow + RPS k, piece l % (PW / 4):
            // one per-lane base and uniform strides (addresses kept out of the loops above)
            constexpr int PPR = PW / 4, RPS = 64 / PPR;
            const int prow = lane / PPR, piece = lane % PPR;
            const bool pok = piece < width / 4;
            const float* tsrc = T + prow * PW + 4 * piece;
            if constexpr (RM) {
#pragma unroll
                for (int k = 0; k < 16 / RPS; ++k) {
                    const int32_t orow = __shfl(rmap, prow + RPS * k);
                    if (pok & (orow >= 0)) {
                        const f32x4 x = *(const f32x4*)(tsrc + RPS * k * PW);
                        __builtin_nontemporal_store(x, (f32x4*)(C + (int64_t)orow * ldc + fh + 4 * piece));
                    }
                }
            } else {
            float* cdst = Crow + (uint32_t)prow * ldc32 + (uint32_t)(fh + 4 * piece);
            const uint32_t cstep = (uint32_t)RPS * ldc32;
#pragma unroll
            for (int k = 0; k < 16 / RPS; ++k) {
                if (pok & (prow + RPS * k < rows_out)) {
                    const f32x4 x = *(const f32x4*)(tsrc + RPS * k * PW);
                    __builtin_nontemporal_store(x, (f32x4*)(cdst + (uint32_t)k * cstep));
                }
            }
            }
        } else if constexpr (RM) {
            int32_t orow[4];
#pragma unroll
            for (int i = 0; i < 4; ++i) orow[i] = __shfl(rmap, 4 * kg + i);
#pragma unroll
            for (int nb = 0; nb < NB; ++nb) {
                const int f = nb * 16 + r;
#pragma unroll
                for (int i = 0; i < 4; ++i) {
                    if ((orow[i] >= 0) & (f < width) & (acc[hf][nb][i] != 0.f))
                        unsafeAtomicAdd(C + (int64_t)orow[i] * ldc + fh + f, acc[hf][nb][i]);
                }
            }
        } else {
            float* adst = Crow + (uint32_t)(4 * kg) * ldc32 + (uint32_t)(fh + r);
#pragma unroll
            for (int nb = 0; nb < NB; ++nb) {
                const int f = nb * 16 + r;
#pragma unroll
                for (int i = 0; i < 4; ++i) {
                    if ((kg * 4 + i < rows_out) & (f < width) & (acc[hf][nb][i] != 0.f))
                        unsafeAtomicAdd(adst + (uint32_t)i * ldc32 + 16 * nb, acc[hf][nb][i]);
                }
            }
        }
        }
    }
}

// one workgroup (one wave) per work item, in the plan's item order (see
// k_tile_routes); 19 KB of LDS per wave at PW = 128, two waves per SIMD
template <int PW, typename DT, int FH = 0, bool RM = false>
__global__ __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(2, 2)))
void k_routes(PlanDev p, const uint16_t* B, int64_t ldb, int K, float* __restrict__ C, int64_t ldc,
              uint32_t mask, int64_t grid, SddArgs sa) {
    // staged rows: 64 (2:4 steps), 32 when only residual steps run (the fused SDDMM)
    constexpr int ROWS = FH > 0 ? 32 : 64;
    __shared__ __align__(16) uint8_t Bs[ROWS * DeepCfg<PW>::RS];
    __shared__ __align__(16) uint8_t Ms[4 * 512];          // record ring: up to 4 slots of 64 x 8 bytes
    static_assert(16 * PW * 4 <= ROWS * DeepCfg<PW>::RS, "the output tile must fit the staging space");
    const int64_t it = p.item_order == 1 ? xcd_window(blockIdx.x, grid) : (int64_t)blockIdx.x;
    if (it < p.n_titems) deep_item<PW, DT, FH, RM>(p, it, B, ldb, K, C, ldc, mask, Bs, Ms, threadIdx.x, sa);
}

// ============================================================ residual route
//
// One wave64 per work item.  The item's entries (the 16 rows of a window, or
// a chunk of one hub row) are contiguous in the residual CSR and are streamed
// as one range: LPR lanes cover an 8*LPR-feature slab of a B row with 16-byte
// loads, the wave's G = 64/LPR lane groups take U consecutive entries each
// (G*U entries and U loads per lane in flight per batch), products are
// accumulated in registers while consecutive entries share a row and flushed
// into the wave's 16-row LDS tile with LDS float atomics when the row
// changes.  The tile is then written once: stored, added to the rows the tile
// kernel wrote, or atomically added for split windows.

__device__ inline void wave_sync() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

template <int LPR, typename DT>
__global__ __launch_bounds__(256) void k_residual(PlanDev p, const uint16_t* __restrict__ B, int64_t ldb, int K,
                                                  float* __restrict__ C, int64_t ldc, uint32_t mask, int64_t grid) {
    constexpr int G = 64 / LPR;
    constexpr int U = 8;
    constexpr int FS = 8 * LPR;                         // features per pass
    __shared__ __align__(16) float tile_all[4][16 * FS];
    __shared__ int32_t rp_all[4][17];
    const int wv = threadIdx.x >> 6, lane = threadIdx.x & 63;
    float* tile = tile_all[wv];
    int32_t* rp = rp_all[wv];
    // blocks (4 items each) are what the dispatcher spreads over the XCDs
    const int64_t it = (p.item_order == 1 ? xcd_window(blockIdx.x, grid) : (int64_t)blockIdx.x) * 4 + wv;
    if (it >= p.n_ritems) return;                       // the whole wave leaves together (no block barriers below)
    const int4 item = p.ritems[it];
    const int64_t w = item.x;
    const int kind = item.y & 0xff;
    const int g = lane / LPR, j = lane % LPR;
    const int fbase = blockIdx.y * FS;
    const int f = fbase + j * 8;
    const bool on = f < K;

    // row boundaries of the item (relative rows 0..15)
    if (lane < 17) {
        if (kind == 2) {
            const int lr = item.y >> 8;
            rp[lane] = lane <= lr ? item.z : item.w;            // all entries belong to row lr
        } else {
            rp[lane] = p.res_rowptr[w * 16 + lane];
        }
    }
    for (int i = lane; i < 16 * FS / 4; i += 64) ((f32x4*)tile)[i] = f32x4{0.f, 0.f, 0.f, 0.f};
    wave_sync();
    const int32_t E0 = rp[0], E1 = (mask & TRSPMM_ROUTE_RESIDUAL) ? rp[16] : rp[0];

    float acc[8] = {0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f};
    int row = 0;
    auto flush = [&]() {
        if (on) {
#pragma unroll
            for (int q = 0; q < 8; ++q) atomicAdd(&tile[row * FS + j * 8 + q], acc[q]);   // ds_add_f32
        }
#pragma unroll
        for (int q = 0; q < 8; ++q) acc[q] = 0.f;
    };
    // All loads are unconditional from clamped in-range addresses and masked
    // afterwards: a predicated load would make the compiler wait for each one.
    const int32_t elast = E1 > E0 ? E1 - 1 : E0;
    const int fl = on ? f : 0;
    auto load_cv = [&](int32_t eb, int32_t (&c)[U], uint16_t (&v)[U]) {
#pragma unroll
        for (int u = 0; u < U; ++u) {
            const int32_t e = eb + u < elast ? eb + u : elast;
            c[u] = p.res_cols[e];
            v[u] = p.res_vals[e];
        }
    };
    // B rows addressed with 32-bit element offsets (the launcher checks ncols * ldb < 2^31)
    const uint16_t* Bf = B + fl;
    const uint32_t ldb32 = (uint32_t)ldb;
    auto load_b = [&](const int32_t (&c)[U], u32x4 (&b)[U]) {
#pragma unroll
        for (int u = 0; u < U; ++u) b[u] = *(const u32x4*)(Bf + (uint32_t)(c[u] > 0 ? c[u] : 0) * ldb32);
    };
    int32_t eb = E0 + g * U;
    if (eb < E1) {
        while (eb >= rp[row + 1]) ++row;                // row of this group's first entry
        int32_t row_end = rp[row + 1];                  // kept in a register; LDS is read on row changes only
        int32_t c0[U], c1[U];
        uint16_t v0[U], v1[U];
        u32x4 b0[U];
        load_cv(eb, c0, v0);
        load_b(c0, b0);
        load_cv(eb + G * U, c1, v1);
#pragma unroll 2
        for (; eb < E1; eb += G * U) {
            // two-stage pipeline: the next batch's gathers and the one after's
            // indices are in flight while this batch is multiplied
            // (loads past the end are clamped to the last entry and never used)
            u32x4 b1[U];
            int32_t c2[U];
            uint16_t v2[U];
            load_b(c1, b1);
            load_cv(eb + 2 * G * U, c2, v2);
#pragma unroll
            for (int u = 0; u < U; ++u) {
                const int32_t e = eb + u;
                if (e >= row_end && e < E1) {
                    flush();
                    do {
                        ++row;
                        row_end = rp[row + 1];
                    } while (e >= row_end);
                }
                // past-the-end / invalid entries read row 0: zero both factors (no 0 * inf)
                const bool ok = e < E1 && c0[u] >= 0;
                const float v = ok ? to_f32<DT>(v0[u]) : 0.f;
                const u16x8 h = __builtin_bit_cast(u16x8, ok ? b0[u] : u32x4{0u, 0u, 0u, 0u});
#pragma unroll
                for (int q = 0; q < 8; ++q) acc[q] = __builtin_fmaf(v, to_f32<DT>(h[q]), acc[q]);
            }
            // the group's next batch starts G*U entries later: usually a new row
            if (eb + G * U < E1 && eb + G * U >= row_end) {
                flush();
                do {
                    ++row;
                    row_end = rp[row + 1];
                } while (eb + G * U >= row_end);
            }
#pragma unroll
            for (int u = 0; u < U; ++u) {
                b0[u] = b1[u];
                c0[u] = c1[u];
                v0[u] = v1[u];
                c1[u] = c2[u];
                v1[u] = v2[u];
            }
        }
        flush();
    }
    wave_sync();

    // write the 16-row tile: 0 = store, 1 = add to the tile kernel's rows, 2 = atomic add
    for (int i = lane; i < 16 * FS / 4; i += 64) {
        const int lr = (4 * i) / FS, ff = (4 * i) % FS;
        const int64_t grow = w * 16 + lr;
        if (grow >= p.m_out || fbase + ff >= K) continue;
        const f32x4 x = ((const f32x4*)tile)[i];
        float* dst = C + grow * ldc + fbase + ff;
        if (kind == 0) {
            __builtin_nontemporal_store(x, (f32x4*)dst);
        } else if (kind == 1) {
            if (rp[lr + 1] > rp[lr]) {
                const f32x4 y = *(const f32x4*)dst;
                *(f32x4*)dst = f32x4{x[0] + y[0], x[1] + y[1], x[2] + y[2], x[3] + y[3]};
            }
        } else {
#pragma unroll
            for (int q = 0; q < 4; ++q)
                if (x[q] != 0.f) unsafeAtomicAdd(dst + q, x[q]);
        }
    }
}

// zero the output rows of split windows (their items accumulate atomically)
__global__ __launch_bounds__(256) void k_zero_rows(const int32_t* __restrict__ split_windows, int64_t m_out, int K,
                                                   float* __restrict__ C, int64_t ldc,
                                                   const int32_t* __restrict__ row_map) {
    const int64_t w = split_windows[blockIdx.x];
    for (int i = threadIdx.x; i < 16 * (K / 4); i += blockDim.x) {
        const int64_t prow = w * 16 + i / (K / 4);
        const int64_t row = row_map ? (int64_t)row_map[prow] : prow;
        const int f = 4 * (i % (K / 4));
        if ((row >= 0) & (row < m_out)) *(f32x4*)(C + row * ldc + f) = f32x4{0.f, 0.f, 0.f, 0.f};
    }
}

// workgroups of a 64-thread kernel the current device holds at once (CUs x
// resident workgroups per CU), remembered per kernel and device; -1 after a
// failed query
template <int F, typename DT>
int64_t resident_blocks() {
    static std::atomic<int64_t> known[64];      // one table per tile-kernel instantiation
    const auto kernel = k_tile_routes<F, DT>;
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess) return -1;
    if (dev >= 0 && dev < 64) {
        const int64_t v = known[dev].load(std::memory_order_relaxed);
        if (v > 0) return v;
    }
    int per_cu = 0, cus = 0;
    hipError_t e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kernel, 64, 0);
    if (e == hipSuccess) e = hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev);
    if (e != hipSuccess || per_cu <= 0 || cus <= 0) {
        set_error(std::string("occupancy query for the persistent tile kernel failed: ") + hipGetErrorString(e));
        return -1;
    }
    const int64_t v = (int64_t)per_cu * cus;
    if (dev >= 0 && dev < 64) known[dev].store(v, std::memory_order_relaxed);
    return v;
}

template <typename DT>
int launch_typed(const PlanDev& p, const void* B, int64_t ldb, int64_t K, float* C, int64_t ldc, uint32_t mask,
                 hipStream_t s) {
    if (p.n_split > 0) {
        hipLaunchKernelGGL(k_zero_rows, dim3((unsigned)p.n_split), dim3(256), 0, s, p.split_windows, p.m_out, (int)K,
                           C, ldc, p.row_map);
        TRSPMM_HIP(hipGetLastError());
    }
    if (p.n_titems > 0) {
        int64_t grid = xcd_grid(p.n_titems);
        if (p.item_order == 2) {
            TRSPMM_CHECK(p.queue != nullptr, "plan has no work queue");
            TRSPMM_CHECK(p.n_titems < 0x7fff0000, "too many work items for the persistent queue");
            const int64_t resident = p.slab == 128 ? resident_blocks<128, DT>() : resident_blocks<64, DT>();
            if (resident < 0) return TRSPMM_EHIP;
            // a multiple of 8, so every XCD's queue starts with workgroups of its own
            grid = std::min<int64_t>(xcd_grid(p.n_titems), std::max<int64_t>(kXcds, resident / kXcds * kXcds));
        }
        TRSPMM_CHECK(grid <= 0x7fffffff, "too many work items");
        // the deep kernel addresses B as a raw buffer of fewer than 2^31 bytes
        const bool deep = p.tile_kernel == 0 && p.item_order != 2 && p.ncols * ldb < ((int64_t)1 << 30);
        if (p.row_map) {
            // a reordered routing: only the deep kernel applies the row map
            if (!deep) {
                set_error("a plan of a reordered routing needs B under 2^30 elements (the deep tile kernel)");
                return TRSPMM_EUNSUPPORTED;
            }
            if (K <= 64)
                hipLaunchKernelGGL((k_routes<64, DT, 0, true>), dim3((unsigned)grid), dim3(64), 0, s, p,
                                   (const uint16_t*)B, ldb, (int)K, C, ldc, mask, grid, SddArgs{});
            else
                hipLaunchKernelGGL((k_routes<128, DT, 0, true>), dim3((unsigned)grid), dim3(64), 0, s, p,
                                   (const uint16_t*)B, ldb, (int)K, C, ldc, mask, grid, SddArgs{});
        } else if (deep) {
            if (K <= 64)
                hipLaunchKernelGGL((k_routes<64, DT>), dim3((unsigned)grid), dim3(64), 0, s, p, (const uint16_t*)B, ldb,
                                   (int)K, C, ldc, mask, grid, SddArgs{});
            else
                hipLaunchKernelGGL((k_routes<128, DT>), dim3((unsigned)grid), dim3(64), 0, s, p, (const uint16_t*)B,
                                   ldb, (int)K, C, ldc, mask, grid, SddArgs{});
        } else if (p.slab == 128)
            hipLaunchKernelGGL((k_tile_routes<128, DT>), dim3((unsigned)grid), dim3(64), 0, s, p, (const uint16_t*)B,
                               ldb, (int)K, C, ldc, mask, grid);
        else
            hipLaunchKernelGGL((k_tile_routes<64, DT>), dim3((unsigned)grid), dim3(64), 0, s, p, (const uint16_t*)B,
                               ldb, (int)K, C, ldc, mask, grid);
        TRSPMM_HIP(hipGetLastError());
    }
    if (p.n_ritems > 0) {
        // lanes per B-row slab: 8 (64 features per pass) or 16 (128 per pass; wider K
        // runs as grid.y slabs so the per-wave LDS tile stays 8 KB)
        const int lpr = K <= 64 ? 8 : 16;
        const int64_t grid = xcd_grid((p.n_ritems + 3) / 4);     // blocks of 4 items
        const unsigned slabs = (unsigned)((K + 8 * lpr - 1) / (8 * lpr));
        TRSPMM_CHECK(grid <= 0x7fffffff, "too many work items");
        const dim3 g((unsigned)grid, slabs);
        if (lpr == 8)
            hipLaunchKernelGGL((k_residual<8, DT>), g, dim3(256), 0, s, p, (const uint16_t*)B, ldb, (int)K, C, ldc,
                               mask, grid);
        else
            hipLaunchKernelGGL((k_residual<16, DT>), g, dim3(256), 0, s, p, (const uint16_t*)B, ldb, (int)K, C, ldc,
                               mask, grid);
        TRSPMM_HIP(hipGetLastError());
    }
    return TRSPMM_OK;
}

template <typename DT>
int launch_fused_typed(const PlanDev& p, const void* dY, int64_t lddy, int64_t K, float* dB, int64_t lddb,
                       const SddArgs& sa, hipStream_t s) {
    if (p.n_split > 0) {
        hipLaunchKernelGGL(k_zero_rows, dim3((unsigned)p.n_split), dim3(256), 0, s, p.split_windows, p.m_out, (int)K,
                           dB, lddb, p.row_map);
        TRSPMM_HIP(hipGetLastError());
    }
    if (p.n_titems == 0) return TRSPMM_OK;
    const int64_t grid = xcd_grid(p.n_titems);
    TRSPMM_CHECK(grid <= 0x7fffffff, "too many work items");
    // K <= 256 is one pass: 128-feature halves, each step's rows gathered per half
    const uint16_t* Y = (const uint16_t*)dY;
    const uint32_t all = (uint32_t)TRSPMM_ROUTE_ALL;
    const dim3 gr((unsigned)grid);
    if (p.row_map) {
        if (K <= 64)
            hipLaunchKernelGGL((k_routes<64, DT, 1, true>), gr, dim3(64), 0, s, p, Y, lddy, (int)K, dB, lddb, all, grid, sa);
        else if (K <= 128)
            hipLaunchKernelGGL((k_routes<128, DT, 1, true>), gr, dim3(64), 0, s, p, Y, lddy, (int)K, dB, lddb, all, grid, sa);
        else
            hipLaunchKernelGGL((k_routes<128, DT, 2, true>), gr, dim3(64), 0, s, p, Y, lddy, (int)K, dB, lddb, all, grid, sa);
    } else if (K <= 64)
        hipLaunchKernelGGL((k_routes<64, DT, 1>), gr, dim3(64), 0, s, p, Y, lddy, (int)K, dB, lddb, all, grid, sa);
    else if (K <= 128)
        hipLaunchKernelGGL((k_routes<128, DT, 1>), gr, dim3(64), 0, s, p, Y, lddy, (int)K, dB, lddb, all, grid, sa);
    else
        hipLaunchKernelGGL((k_routes<128, DT, 2>), gr, dim3(64), 0, s, p, Y, lddy, (int)K, dB, lddb, all, grid, sa);
    TRSPMM_HIP(hipGetLastError());
    return TRSPMM_OK;
}

}  // namespace

int fused_sddmm_supported(const PlanDev& p) {
    return p.n_tiles == 0 && p.res_mfma && p.tile_kernel == 0 && p.item_order != 2;
}

int launch_spmm_sddmm(const PlanDev& p, const void* dY, int64_t lddy, int64_t K, float* dB, int64_t lddb,
                      const void* X, int64_t ldx, int64_t x_rows, float* D, hipStream_t s) {
    TRSPMM_CHECK(fused_sddmm_supported(p), "fused SDDMM: the plan must hold TC sub-blocks and residual records only "
                                           "(no 2:4 tiles), on the matrix-core engine, with the default tile kernel "
                                           "and item order");
    TRSPMM_CHECK(K > 0 && K % 8 == 0 && K <= (int64_t)1 << 24, "K must be a positive multiple of 8");
    TRSPMM_CHECK(lddy >= K && lddy % 8 == 0 && ldx >= K && ldx % 8 == 0, "dY and X strides must be >= K, multiples of 8");
    TRSPMM_CHECK(lddb >= K && lddb % 4 == 0, "dB stride must be >= K and a multiple of 4");
    TRSPMM_CHECK(((uintptr_t)dY & 15) == 0 && ((uintptr_t)dB & 15) == 0 && ((uintptr_t)X & 15) == 0,
                 "dY, dB and X must be 16-byte aligned");
    TRSPMM_CHECK(p.ncols * lddy < ((int64_t)1 << 30) && x_rows * ldx < ((int64_t)1 << 30),
                 "fused SDDMM: dY and X must hold fewer than 2^30 elements");
    TRSPMM_CHECK(p.n_rval + p.n_tval == 0 || D != nullptr, "fused SDDMM: D must not be null");
    if (p.windows == 0) return TRSPMM_OK;
    const SddArgs sa{(const uint16_t*)X, ldx, x_rows, D};
    if (p.dtype == TRSPMM_BF16) return launch_fused_typed<BF16>(p, dY, lddy, K, dB, lddb, sa, s);
    return launch_fused_typed<F16>(p, dY, lddy, K, dB, lddb, sa, s);
}

int launch_spmm(const PlanDev& p, const void* B, int64_t ldb, int64_t K, float* C, int64_t ldc, uint32_t mask,
                hipStream_t s) {
    TRSPMM_CHECK(K > 0 && K % 8 == 0, "K must be a positive multiple of 8");
    TRSPMM_CHECK(ldb >= K && ldb % 8 == 0, "ldb must be >= K and a multiple of 8");
    TRSPMM_CHECK(ldc >= K && ldc % 4 == 0, "ldc must be >= K and a multiple of 4");
    TRSPMM_CHECK(((uintptr_t)B & 15) == 0 && ((uintptr_t)C & 15) == 0, "B and C must be 16-byte aligned");
    TRSPMM_CHECK(K <= (int64_t)1 << 24, "K too large");
    TRSPMM_CHECK(p.ncols * ldb < ((int64_t)1 << 31), "B must hold fewer than 2^31 elements");
    if (p.windows == 0) return TRSPMM_OK;
    if (p.dtype == TRSPMM_BF16) return launch_typed<BF16>(p, B, ldb, K, C, ldc, mask, s);
    return launch_typed<F16>(p, B, ldb, K, C, ldc, mask, s);
}

}  // namespace trspmm
