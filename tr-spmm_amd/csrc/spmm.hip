// Fused three-route SpMM for gfx950: one wave64 workgroup owns one 16-row
// window and one feature slab, runs the window's dense-tile route on MFMA,
// its 2:4 route on SMFMAC and its residual route on the vector ALUs, and
// writes each output row exactly once.
//
// Reference semantics (TR-source/SpMM/mGCNkernel_online.cu):
//   TC route     :54-241  C[16w+r, f] += sum_q A_tc[r, q] * B[t_column[8s+q], f]
//   2:4 route    :812-933 C[16w+r, f] += sum over the tile's 16 columns
//   residual     mGCNkernel.cu:977-1171 with spmm_utils/compute.h:53-87
//   add kernel   :1314-1324 (not needed: the routes share one accumulator)
// The reference accumulates TC and residual products in fp16 and writes the
// 2:4 route to a second buffer; here all routes accumulate in fp32.
//
// Per slab of F features (F = 64 here) the workgroup:
//   1. TC: for every 4 sub-blocks (32 columns) gathers the 32 B-row segments
//      into LDS, decodes its A fragment from the occupancy bits (lane l holds
//      row l%16 of sub-block l/16: 8 values located by popcount), and issues
//      one v_mfma_f32_16x16x32 per 16 features.
//   2. 2:4: for every 4 tiles (64 columns) gathers 64 B-row segments into LDS,
//      loads its compressed A row (16 bytes) and 2-bit index word, and issues
//      one v_smfmac_f32_16x16x64 per 16 features.
//   3. stores the MFMA accumulators into an LDS C tile, then the residual
//      parts run on groups of F/8 lanes (16-byte B loads per lane, fp32 FMA)
//      and add into the C tile with LDS float atomics.
//   4. streams the C tile to HBM with non-temporal 16-byte stores.
#include "plan.h"

namespace trspmm {

namespace {

typedef uint16_t u16x8 __attribute__((ext_vector_type(8)));
typedef uint16_t u16x16 __attribute__((ext_vector_type(16)));
typedef short s16x4 __attribute__((__vector_size__(8)));
typedef short s16x8 __attribute__((ext_vector_type(8)));
typedef short s16x16 __attribute__((ext_vector_type(16)));

// LDS transpose read (gfx950 ds_read_b64_tr_b16).  Within each group of 16
// lanes, lane L supplies the address of 4 consecutive 16-bit values (chunk_L)
// and lane j receives element i = chunk_{4i + j/4}[j % 4].  With row-major
// staged B rows, lane L pointing at row (base + L/4), features
// (f + 4*(L%4) .. +4), lane j receives rows base..base+3 of feature f + j:
// four consecutive K entries of one MFMA B-operand column in one instruction.
__device__ inline s16x4 lds_tr16(const uint16_t* p) {
    return __builtin_amdgcn_ds_read_tr16_b64_v4i16((__attribute__((address_space(3))) s16x4*)p);
}

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

// Gathers rows cols[0..nrows) (segment [f0, f0+fw) of each) into Bs as
// R rows of F 16-bit values; rows >= nrows, columns < 0 and features >= fw
// are zero-filled.  Row stride LDS_STRIDE keeps the 4 lane groups of an MFMA
// operand read on disjoint LDS banks.
template <int F, int R, int LDS_STRIDE>
__device__ inline void stage_rows(uint16_t* __restrict__ Bs, const uint16_t* __restrict__ B, int64_t ldb,
                                  int f0, int fw, const int32_t* __restrict__ cols, int nrows, int lane) {
    constexpr int PIECES = F / 8;             // 16-byte pieces per row segment
    constexpr int PER_LANE = R * PIECES / 64;
#pragma unroll
    for (int k = 0; k < PER_LANE; ++k) {
        const int c = lane + 64 * k;
        const int row = c / PIECES, piece = c % PIECES;
        const int col = row < nrows ? cols[row] : -1;
        u32x4 v = {0u, 0u, 0u, 0u};
        if (col >= 0 && piece * 8 < fw)
            v = *(const u32x4*)(B + (int64_t)col * ldb + f0 + piece * 8);   // B rows stay cached: reused across windows
        u32x2* dst = (u32x2*)(Bs + row * LDS_STRIDE + piece * 8);
        dst[0] = u32x2{v.x, v.y};
        dst[1] = u32x2{v.z, v.w};
    }
}

template <int F, typename DT>
__global__ __launch_bounds__(64) void k_routed_spmm(PlanDev p, const uint16_t* __restrict__ B, int64_t ldb, int K,
                                                    float* __restrict__ C, int64_t ldc, uint32_t mask, int64_t grid) {
    using X8 = typename DT::x8;
    using X16 = typename DT::x16;
    constexpr int LDS_STRIDE = F + 4;
    constexpr int NB = F / 16;
    __shared__ __align__(16) uint16_t Bs[64 * LDS_STRIDE];
    __shared__ __align__(16) float Cs[16 * F];

    const int lane = threadIdx.x;
    const int64_t w = xcd_window(blockIdx.x, grid);
    if (w >= p.windows) return;
    const int64_t row0 = w * 16;
    const int32_t s0 = (mask & TRSPMM_ROUTE_TC) && p.n_tsub ? p.tc_sub_off[w] : 0;
    const int32_t s1 = (mask & TRSPMM_ROUTE_TC) && p.n_tsub ? p.tc_sub_off[w + 1] : 0;
    const int32_t t0 = (mask & TRSPMM_ROUTE_SPTC) && p.n_tiles ? p.sp_tile_off[w] : 0;
    const int32_t t1 = (mask & TRSPMM_ROUTE_SPTC) && p.n_tiles ? p.sp_tile_off[w + 1] : 0;
    const int32_t q0 = (mask & TRSPMM_ROUTE_RESIDUAL) && p.n_rparts ? p.res_off[w] : 0;
    const int32_t q1 = (mask & TRSPMM_ROUTE_RESIDUAL) && p.n_rparts ? p.res_off[w + 1] : 0;
    const int r = lane & 15, kg = lane >> 4;

    for (int f0 = 0; f0 < K; f0 += F) {
        const int fw = K - f0 < F ? K - f0 : F;
        f32x4 acc[NB];
#pragma unroll
        for (int nb = 0; nb < NB; ++nb) acc[nb] = f32x4{0.f, 0.f, 0.f, 0.f};

        // ---------------- dense 16x8 sub-blocks, 4 per MFMA (K = 32 columns)
        for (int32_t sb = s0; sb < s1; sb += 4) {
            __syncthreads();
            const int nvalid = (s1 - sb) * 8 < 32 ? (s1 - sb) * 8 : 32;
            stage_rows<F, 32, LDS_STRIDE>(Bs, B, ldb, f0, fw, p.tc_cols + (int64_t)sb * 8, nvalid, lane);
            u16x8 abits = {0, 0, 0, 0, 0, 0, 0, 0};
            const int32_t my = sb + kg;
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
            const X8 a = __builtin_bit_cast(X8, abits);
            __syncthreads();
            // B operand of v_mfma_f32_16x16x32: lane (kg, n) holds staged rows 8kg..8kg+7 of feature n
            const uint16_t* bp = Bs + (kg * 8 + (r >> 2)) * LDS_STRIDE + 4 * (r & 3);
#pragma unroll
            for (int nb = 0; nb < NB; ++nb) {
                const s16x4 b0 = lds_tr16(bp + nb * 16), b1 = lds_tr16(bp + 4 * LDS_STRIDE + nb * 16);
                const s16x8 bb = {b0[0], b0[1], b0[2], b0[3], b1[0], b1[1], b1[2], b1[3]};
                acc[nb] = mma_dense<DT>(a, __builtin_bit_cast(X8, bb), acc[nb]);
            }
        }

        // ---------------- 2:4 tiles, 4 per SMFMAC (K = 64 columns)
        for (int32_t t = t0; t < t1; t += 4) {
            __syncthreads();
            const int nvalid = (t1 - t) * 16 < 64 ? (t1 - t) * 16 : 64;
            stage_rows<F, 64, LDS_STRIDE>(Bs, B, ldb, f0, fw, p.sp_cols + (int64_t)t * 16, nvalid, lane);
            u32x4 av = {0u, 0u, 0u, 0u};
            int idx = 0x4444;                   // (0,1) in every group: valid and multiplied by zeros
            const int32_t my = t + kg;
            if (my < t1) {
                av = *(const u32x4*)(p.sp_a + ((int64_t)my * 16 + r) * 4);
                idx = p.sp_idx[(int64_t)my * 16 + r];
            }
            const X8 a = __builtin_bit_cast(X8, av);
            __syncthreads();
            // B operand of v_smfmac_f32_16x16x64 (measured lane layout): slot q of lane
            // (kg, n) holds staged row 32*(q/8) + 8*kg + q%8 of feature n, where staged
            // row 16t + 4g + p is column p of group g of the t-th tile (the A layout:
            // lane (kg, r) = tile kg, row r, groups 0..3, nibble g = p0 | p1 << 2).
            const uint16_t* bp = Bs + (kg * 8 + (r >> 2)) * LDS_STRIDE + 4 * (r & 3);
#pragma unroll
            for (int nb = 0; nb < NB; ++nb) {
                const s16x4 b0 = lds_tr16(bp + nb * 16), b1 = lds_tr16(bp + 4 * LDS_STRIDE + nb * 16);
                const s16x4 b2 = lds_tr16(bp + 32 * LDS_STRIDE + nb * 16);
                const s16x4 b3 = lds_tr16(bp + 36 * LDS_STRIDE + nb * 16);
                const s16x16 bb = {b0[0], b0[1], b0[2], b0[3], b1[0], b1[1], b1[2], b1[3],
                                   b2[0], b2[1], b2[2], b2[3], b3[0], b3[1], b3[2], b3[3]};
                acc[nb] = mma_sparse<DT>(a, __builtin_bit_cast(X16, bb), acc[nb], idx);
            }
        }

        // ---------------- accumulators -> LDS C tile (lane holds rows 4kg..4kg+3, feature r)
        __syncthreads();
#pragma unroll
        for (int nb = 0; nb < NB; ++nb)
#pragma unroll
            for (int i = 0; i < 4; ++i) Cs[(kg * 4 + i) * F + nb * 16 + r] = acc[nb][i];
        __syncthreads();

        // ---------------- residual parts: F/8 lanes per B-row segment
        if (q1 > q0) {
            constexpr int LPR = F / 8;
            constexpr int GROUPS = 64 / LPR;
            const int g = lane / LPR, j = lane % LPR;
            const bool on = j * 8 < fw;
            for (int32_t pi = q0 + g; pi < q1; pi += GROUPS) {
                const int32_t row = p.res_row[pi], beg = p.res_beg[pi], cnt = p.res_cnt[pi];
                float x[8] = {0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f};
                if (on) {
                    for (int32_t e = beg; e < beg + cnt; ++e) {
                        const int32_t col = p.res_cols[e];
                        if (col < 0) continue;
                        const float v = to_f32<DT>(p.res_vals[e]);
                        const u16x8 bv = __builtin_bit_cast(
                            u16x8, *(const u32x4*)(B + (int64_t)col * ldb + f0 + j * 8));
#pragma unroll
                        for (int q = 0; q < 8; ++q) x[q] = __builtin_fmaf(v, to_f32<DT>(bv[q]), x[q]);
                    }
                }
                const int64_t lr = row - row0;
                if (on && lr >= 0 && lr < 16) {
#pragma unroll
                    for (int q = 0; q < 8; ++q) atomicAdd(&Cs[lr * F + j * 8 + q], x[q]);
                }
            }
            __syncthreads();
        }

        // ---------------- C tile -> HBM (each output row written once)
#pragma unroll
        for (int i = lane; i < 16 * F / 4; i += 64) {
            const int rr = (i * 4) / F, f = (i * 4) % F;
            const int64_t row = row0 + rr;
            if (row < p.m_out && f < fw)
                __builtin_nontemporal_store(*(const f32x4*)(Cs + rr * F + f), (f32x4*)(C + row * ldc + f0 + f));
        }
        __syncthreads();
    }
}

}  // namespace

int launch_spmm(const PlanDev& p, const void* B, int64_t ldb, int64_t K, float* C, int64_t ldc, uint32_t mask,
                hipStream_t s) {
    TRSPMM_CHECK(K > 0 && K % 8 == 0, "K must be a positive multiple of 8");
    TRSPMM_CHECK(ldb >= K && ldb % 8 == 0, "ldb must be >= K and a multiple of 8");
    TRSPMM_CHECK(ldc >= K && ldc % 4 == 0, "ldc must be >= K and a multiple of 4");
    TRSPMM_CHECK(((uintptr_t)B & 15) == 0 && ((uintptr_t)C & 15) == 0, "B and C must be 16-byte aligned");
    TRSPMM_CHECK(K <= (int64_t)1 << 30, "K too large");
    if (p.windows == 0) return TRSPMM_OK;
    const int64_t grid = xcd_grid(p.windows);
    TRSPMM_CHECK(grid <= 0x7fffffff, "too many windows");
    if (p.dtype == TRSPMM_BF16)
        hipLaunchKernelGGL((k_routed_spmm<64, BF16>), dim3((unsigned)grid), dim3(64), 0, s, p, (const uint16_t*)B,
                           ldb, (int)K, C, ldc, mask, grid);
    else
        hipLaunchKernelGGL((k_routed_spmm<64, F16>), dim3((unsigned)grid), dim3(64), 0, s, p, (const uint16_t*)B,
                           ldb, (int)K, C, ldc, mask, grid);
    TRSPMM_HIP(hipGetLastError());
    return TRSPMM_OK;
}

}  // namespace trspmm
