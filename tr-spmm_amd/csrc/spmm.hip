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
//   k_residual    one wave64 per row window: each row's residual entries are
//                 split over the lane groups, 16-byte B loads per lane (8 per
//                 lane in flight), fp32 FMA, cross-group butterfly reduction
//                 with lane shuffles; stores the rows (or adds to rows the
//                 tile kernel wrote).
// Every output element is written by exactly one work item except in split
// windows, whose rows are zeroed and accumulated with fp32 atomics.
#include <cstdlib>

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
// rows) or 4 2:4 tiles (64 gathered rows).  Lane l gathers the 16-byte pieces
// c = l + 64k of the step's rows (row c / (F/8), piece c % (F/8)).
template <int F>
struct Gather {
    static constexpr int PIECES = F / 8;
    static constexpr int MAXK = 64 * PIECES / 64;   // pieces per lane for a 64-row step
    u32x4 v[MAXK];
};

// cols[row * STRIDE] is the column of staged row `row` (STRIDE 2: the low word
// of a residual entry record)
template <int F, int R, int STRIDE = 1>
__device__ inline void gather_rows(Gather<F>& g, const uint16_t* __restrict__ B, int64_t ldb, int f0, int fw,
                                   const int32_t* __restrict__ cols, int nrows, int lane) {
    constexpr int PIECES = F / 8;
    // unconditional loads from clamped in-range addresses, masked afterwards
    // (a predicated load would make the compiler wait for each one)
    int col[R * PIECES / 64];
#pragma unroll
    for (int k = 0; k < R * PIECES / 64; ++k) {
        const int row = (lane + 64 * k) / PIECES;
        col[k] = cols[(row < nrows ? row : nrows - 1) * STRIDE];
    }
#pragma unroll
    for (int k = 0; k < R * PIECES / 64; ++k) {
        const int c = lane + 64 * k;
        const int row = c / PIECES, piece = c % PIECES;
        const bool ok = row < nrows && col[k] >= 0 && piece * 8 < fw;
        const u32x4 x = *(const u32x4*)(B + (int64_t)(ok ? col[k] : 0) * ldb + f0 + (ok ? piece * 8 : 0));
        g.v[k] = ok ? x : u32x4{0u, 0u, 0u, 0u};
    }
}

template <int F, int R, int LDS_STRIDE>
__device__ inline void park_rows(const Gather<F>& g, uint16_t* __restrict__ Bs, int lane) {
    constexpr int PIECES = F / 8;
#pragma unroll
    for (int k = 0; k < R * PIECES / 64; ++k) {
        const int c = lane + 64 * k;
        const int row = c / PIECES, piece = c % PIECES;
        u32x2* dst = (u32x2*)(Bs + row * LDS_STRIDE + piece * 8);
        dst[0] = u32x2{g.v[k].x, g.v[k].y};
        dst[1] = u32x2{g.v[k].z, g.v[k].w};
    }
}

// A fragment of a TC step: lane (kg, r) holds row r of sub-block sb+kg, the
// 8 slots decoded from the occupancy bits (values stored in bit order).
__device__ inline u16x8 tc_fragment(const PlanDev& p, int32_t my, int32_t s1, int r) {
    u16x8 abits = {0, 0, 0, 0, 0, 0, 0, 0};
    if (my < s1) {
        const uint64_t lo = p.tc_bits[2 * (int64_t)my], hi = p.tc_bits[2 * (int64_t)my + 1];
        const uint64_t word = r < 8 ? lo : hi;
        const int sh = (r & 7) * 8;
        const uint32_t bits = (uint32_t)(word >> sh) & 0xffu;
        int vi = p.tc_valoff[my] + (r < 8 ? 0 : __popcll(lo)) + __popcll(word & ((1ull << sh) - 1ull));
#pragma unroll
        for (int q = 0; q < 8; ++q)
            if ((bits >> q) & 1u) abits[q] = p.tc_vals[vi++];
    }
    return abits;
}

template <int F, typename DT>
__global__ __launch_bounds__(64) void k_tile_routes(PlanDev p, const uint16_t* __restrict__ B, int64_t ldb, int K,
                                                    float* __restrict__ C, int64_t ldc, uint32_t mask, int64_t grid) {
    using X8 = typename DT::x8;
    using X16 = typename DT::x16;
    constexpr int LDS_STRIDE = F + 4;     // keeps the 4 lane groups of an operand read on disjoint banks
    constexpr int NB = F / 16;
    __shared__ __align__(16) uint16_t Bs[64 * LDS_STRIDE];

    const int lane = threadIdx.x;
    const int64_t it = xcd_window(blockIdx.x, grid);
    if (it >= p.n_titems) return;
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
    const int nsteps = nts + ntt + nte;
    const int64_t row0 = w * 16;
    const uint16_t* bp = Bs + (kg * 8 + (r >> 2)) * LDS_STRIDE + 4 * (r & 3);

    for (int f0 = 0; f0 < K; f0 += F) {
        const int fw = K - f0 < F ? K - f0 : F;
        f32x4 acc[NB];
#pragma unroll
        for (int nb = 0; nb < NB; ++nb) acc[nb] = f32x4{0.f, 0.f, 0.f, 0.f};

        Gather<F> g;
        const int sp0 = nts, re0 = nts + ntt;      // first 2:4 step, first residual step
        auto issue = [&](int step) {
            if (step < sp0) {
                const int32_t sb = s0 + 4 * step;
                gather_rows<F, 32>(g, B, ldb, f0, fw, p.tc_cols + (int64_t)sb * 8, (s1 - sb) * 8 < 32 ? (s1 - sb) * 8 : 32,
                                   lane);
            } else if (step < re0) {
                const int32_t t = t0 + 4 * (step - sp0);
                gather_rows<F, 64>(g, B, ldb, f0, fw, p.sp_cols + (int64_t)t * 16, (t1 - t) * 16 < 64 ? (t1 - t) * 16 : 64,
                                   lane);
            } else if (step < nsteps) {
                const int32_t eb = e0 + 32 * (step - re0);
                gather_rows<F, 32, 2>(g, B, ldb, f0, fw, (const int32_t*)(p.res_rec + eb), e1 - eb < 32 ? e1 - eb : 32,
                                      lane);
            }
        };
        issue(0);
        for (int step = 0; step < nsteps; ++step) {
            __syncthreads();                       // previous step's LDS reads are done
            if (step < sp0 || step >= re0) park_rows<F, 32, LDS_STRIDE>(g, Bs, lane);
            else park_rows<F, 64, LDS_STRIDE>(g, Bs, lane);
            issue(step + 1);                       // next step's gathers overlap this step's math
            if (step < sp0 || step >= re0) {
                u16x8 abits;
                if (step < sp0) {
                    abits = tc_fragment(p, s0 + 4 * step + kg, s1, r);
                } else {
                    // residual entries as a 16x32 one-hot fragment: staged row k = entry k, and
                    // A[r][k] = value of entry k if its local row is r (records: col | val << 32 | lrow << 48)
                    const int32_t eb = e0 + 32 * (step - re0) + 8 * kg;
#pragma unroll
                    for (int q = 0; q < 8; ++q) {
                        const int32_t e = eb + q < e1 ? eb + q : e1 - 1;
                        const uint64_t rec = p.res_rec[e];
                        const bool hit = eb + q < e1 && (int)((rec >> 48) & 0xffu) == r && (int32_t)rec >= 0;
                        abits[q] = hit ? (uint16_t)(rec >> 32) : (uint16_t)0;
                    }
                }
                const X8 a = __builtin_bit_cast(X8, abits);
                __syncthreads();
#pragma unroll
                for (int nb = 0; nb < NB; ++nb) {
                    // B operand of v_mfma_f32_16x16x32: lane (kg, n) holds staged rows 8kg..8kg+7, feature n
                    const s16x4 b0 = lds_tr16(bp + nb * 16), b1 = lds_tr16(bp + 4 * LDS_STRIDE + nb * 16);
                    const s16x8 bb = {b0[0], b0[1], b0[2], b0[3], b1[0], b1[1], b1[2], b1[3]};
                    acc[nb] = mma_dense<DT>(a, __builtin_bit_cast(X8, bb), acc[nb]);
                }
            } else {
                const int32_t my = t0 + 4 * (step - sp0) + kg;
                u32x4 av = {0u, 0u, 0u, 0u};
                int idx = 0x4444;                  // (0,1) in every group: valid, multiplied by zeros
                if (my < t1) {
                    av = *(const u32x4*)(p.sp_a + ((int64_t)my * 16 + r) * 4);
                    idx = p.sp_idx[(int64_t)my * 16 + r];
                }
                const X8 a = __builtin_bit_cast(X8, av);
                __syncthreads();
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
                    acc[nb] = mma_sparse<DT>(a, __builtin_bit_cast(X16, bb), acc[nb], idx);
                }
            }
        }
        // accumulator element i of lane (kg, n) = C[16w + 4kg + i, f0 + 16nb + n]
#pragma unroll
        for (int nb = 0; nb < NB; ++nb) {
            const int f = nb * 16 + r;
#pragma unroll
            for (int i = 0; i < 4; ++i) {
                const int64_t row = row0 + kg * 4 + i;
                if (row < p.m_out && f < fw) {
                    float* dst = C + row * ldc + f0 + f;
                    if (!split) __builtin_nontemporal_store(acc[nb][i], dst);
                    else if (acc[nb][i] != 0.f) unsafeAtomicAdd(dst, acc[nb][i]);
                }
            }
        }
    }
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
    const int64_t it = xcd_window(blockIdx.x, grid) * 4 + wv;
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
                                                   float* __restrict__ C, int64_t ldc) {
    const int64_t w = split_windows[blockIdx.x];
    for (int i = threadIdx.x; i < 16 * (K / 4); i += blockDim.x) {
        const int64_t row = w * 16 + i / (K / 4);
        const int f = 4 * (i % (K / 4));
        if (row < m_out) *(f32x4*)(C + row * ldc + f) = f32x4{0.f, 0.f, 0.f, 0.f};
    }
}

template <typename DT>
int launch_typed(const PlanDev& p, const void* B, int64_t ldb, int64_t K, float* C, int64_t ldc, uint32_t mask,
                 hipStream_t s) {
    if (p.n_split > 0) {
        hipLaunchKernelGGL(k_zero_rows, dim3((unsigned)p.n_split), dim3(256), 0, s, p.split_windows, p.m_out, (int)K,
                           C, ldc);
        TRSPMM_HIP(hipGetLastError());
    }
    if (p.n_titems > 0) {
        const int64_t grid = xcd_grid(p.n_titems);
        TRSPMM_CHECK(grid <= 0x7fffffff, "too many work items");
        if (p.slab == 128)
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

}  // namespace

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
