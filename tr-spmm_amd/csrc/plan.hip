// Device plan builder: reference-format routed metadata -> gfx950 lane layouts.
//
// Input contract (the op's inputs, TR-source/SpMM/mGCN_online.cpp:760-932):
// TC parts carry t_window_row = 16*window (block_online.cpp:493) and are in
// window order; residual parts (long and short lists) are in row order; 2:4
// groups are in window order with s_offsets[W+1].  The plan keeps the same
// nonzeros and values; only layouts change:
//   * per-window ranges for every route (so one workgroup owns one window),
//   * 2:4 tiles repacked so a lane reads its row's 4 groups as 16 contiguous
//     bytes plus one 16-bit index word (the SMFMAC 16x16x64 A-operand form),
//   * long + short residual parts merged into one per-window part list,
//   * values converted once to the plan's 16-bit type (fp16 or bf16),
//   * columns outside [0, ncols) replaced by -1 (contribute nothing), the
//     bounds rule the reference kernels apply (mGCNkernel_online.cu:296-306).
#include "plan.h"
#include "scan.h"

namespace trspmm {

namespace {

__device__ inline int32_t clean_col(int32_t c, int64_t ncols) {
    return (c >= 0 && (int64_t)c < ncols) ? c : -1;
}

__device__ inline uint16_t to_plan_bits(const void* src, int64_t i, int src_is_f32, int dtype) {
    if (src_is_f32) {
        const float f = ((const float*)src)[i];
        return dtype == TRSPMM_BF16 ? f32_to_bf16_bits(f) : f32_to_f16_bits(f);
    }
    const uint16_t h = ((const uint16_t*)src)[i];           // fp16 input
    if (dtype == TRSPMM_F16) return h;
    return f32_to_bf16_bits((float)__builtin_bit_cast(_Float16, h));
}

// per window w in [0, W]: the first TC part / residual part at or after w
__global__ void k_window_bounds(const int32_t* __restrict__ t_window_row, int64_t n_tparts,
                                const int32_t* __restrict__ t_row_offset,
                                const int32_t* __restrict__ c_row, int64_t n_cparts,
                                const int32_t* __restrict__ cs_row, int64_t n_csparts, int64_t W,
                                int32_t* __restrict__ tc_sub_off, int32_t* __restrict__ lb_long,
                                int32_t* __restrict__ lb_short) {
    const int64_t w = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (w > W) return;
    const int32_t row = (int32_t)(w * 16);
    const int64_t q = lower_bound(t_window_row, n_tparts, row);
    tc_sub_off[w] = n_tparts ? t_row_offset[q] : 0;
    lb_long[w] = (int32_t)lower_bound(c_row, n_cparts, row);
    lb_short[w] = (int32_t)lower_bound(cs_row, n_csparts, row);
}

__global__ void k_residual_parts(const int32_t* __restrict__ lb_long, const int32_t* __restrict__ lb_short,
                                 const int32_t* __restrict__ c_row_offset, const int32_t* __restrict__ c_row,
                                 const int32_t* __restrict__ cs_row_offset, const int32_t* __restrict__ cs_row,
                                 int64_t n_cval, int64_t W, int32_t* __restrict__ res_off,
                                 int32_t* __restrict__ res_row, int32_t* __restrict__ res_beg,
                                 int32_t* __restrict__ res_cnt) {
    const int64_t w = blockIdx.x;
    if (w >= W) return;
    const int32_t l0 = lb_long[w], l1 = lb_long[w + 1], s0 = lb_short[w], s1 = lb_short[w + 1];
    const int32_t base = l0 + s0;
    if (threadIdx.x == 0) {
        res_off[w] = base;
        if (w == W - 1) res_off[W] = l1 + s1;
    }
    for (int32_t i = threadIdx.x; i < (l1 - l0) + (s1 - s0); i += blockDim.x) {
        const int32_t u = base + i;
        if (i < l1 - l0) {
            const int32_t p = l0 + i;
            res_row[u] = c_row[p];
            res_beg[u] = c_row_offset[p];
            res_cnt[u] = c_row_offset[p + 1] - c_row_offset[p];
        } else {
            const int32_t p = s0 + (i - (l1 - l0));
            res_row[u] = cs_row[p];
            res_beg[u] = (int32_t)n_cval + cs_row_offset[p];
            res_cnt[u] = cs_row_offset[p + 1] - cs_row_offset[p];
        }
    }
}

__global__ void k_copy_cols(const int32_t* __restrict__ src, int64_t n, int64_t ncols, int32_t* __restrict__ dst) {
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
        dst[i] = clean_col(src[i], ncols);
}

__global__ void k_copy_vals(const void* __restrict__ src, int64_t n, int src_is_f32, int dtype,
                            uint16_t* __restrict__ dst) {
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
        dst[i] = to_plan_bits(src, i, src_is_f32, dtype);
}

__global__ void k_tile_counts(const int32_t* __restrict__ s_offsets, int64_t W, int32_t* __restrict__ cnt) {
    const int64_t w = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (w < W) cnt[w] = (s_offsets[w + 1] - s_offsets[w] + 3) / 4;
}

// One workgroup per window; thread (tile j, row r, group c) fills one A word.
// Slot normalisation follows block_online.cpp:190-210: both hits -> (p0<p1);
// one hit at q -> a zero-valued companion index (1 if q == 0 else 0); no hit
// or a missing group -> indices (0, 1) with zero values.
__global__ void k_tiles_from_groups(const int32_t* __restrict__ s_offsets, const int32_t* __restrict__ s_column,
                                    const float* __restrict__ s_value, const int8_t* __restrict__ s_pos,
                                    const int32_t* __restrict__ tile_off, int64_t W, int window,
                                    int64_t ncols, int dtype, int32_t* __restrict__ sp_cols,
                                    uint32_t* __restrict__ sp_a, uint16_t* __restrict__ sp_idx) {
    const int64_t w = blockIdx.x;
    if (w >= W) return;
    const int32_t g0 = s_offsets[w], g1 = s_offsets[w + 1];
    const int32_t t0 = tile_off[w], nt = tile_off[w + 1] - t0;
    for (int32_t i = threadIdx.x; i < nt * 64; i += blockDim.x) {
        const int32_t j = i >> 6, r = (i >> 2) & 15, c = i & 3;
        const int64_t t = t0 + j;
        const int32_t g = g0 + 4 * j + c;
        float v0 = 0.f, v1 = 0.f;
        int p0 = 0, p1 = 1;
        if (g < g1 && r < window) {
            const int64_t b = ((int64_t)g * window + r) * 2;
            const int q0 = s_pos[b], q1 = s_pos[b + 1];
            const float x0 = s_value[b], x1 = s_value[b + 1];
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
        const uint32_t lo = dtype == TRSPMM_BF16 ? f32_to_bf16_bits(v0) : f32_to_f16_bits(v0);
        const uint32_t hi = dtype == TRSPMM_BF16 ? f32_to_bf16_bits(v1) : f32_to_f16_bits(v1);
        sp_a[(t * 16 + r) * 4 + c] = lo | (hi << 16);
        // the four groups of one row share one index word: thread c == 0 writes it
        if (c == 0) {
            uint32_t idx = 0;
            for (int cc = 0; cc < 4; ++cc) {
                const int32_t gg = g0 + 4 * j + cc;
                int a0 = 0, a1 = 1;
                if (gg < g1 && r < window) {
                    const int64_t b = ((int64_t)gg * window + r) * 2;
                    const int q0 = s_pos[b], q1 = s_pos[b + 1];
                    if (q0 >= 0 && q1 >= 0) { a0 = q0 < q1 ? q0 : q1; a1 = q0 < q1 ? q1 : q0; }
                    else if (q0 >= 0 || q1 >= 0) {
                        const int q = q0 >= 0 ? q0 : q1;
                        const int dummy = q == 0 ? 1 : 0;
                        a0 = dummy < q ? dummy : q;
                        a1 = dummy < q ? q : dummy;
                    }
                }
                idx |= (uint32_t)((a0 & 3) | ((a1 & 3) << 2)) << (4 * cc);
            }
            sp_idx[t * 16 + r] = (uint16_t)idx;
        }
        if (r == 0) {
            for (int q = 0; q < 4; ++q)
                sp_cols[t * 16 + 4 * c + q] = g < g1 ? clean_col(s_column[4 * g + q], ncols) : -1;
        }
    }
}

// mma.sp packed tiles (block_online.cpp:796-841 layout) -> plan tiles:
// row r of tile t lives at s_packed_a[t][r % 8][c][r / 8] and in the nibbles
// of s_packed_meta[t][r % 8] starting at bit 16 * (r / 8).
__global__ void k_tiles_from_packed(const int32_t* __restrict__ s_tile_column, const uint32_t* __restrict__ s_packed_a,
                                    const uint32_t* __restrict__ s_packed_meta, int64_t T, int64_t ncols,
                                    int32_t* __restrict__ sp_cols, uint32_t* __restrict__ sp_a,
                                    uint16_t* __restrict__ sp_idx) {
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < T * 64; i += (int64_t)gridDim.x * blockDim.x) {
        const int64_t t = i >> 6;
        const int r = (int)((i >> 2) & 15), c = (int)(i & 3);
        sp_a[(t * 16 + r) * 4 + c] = s_packed_a[((t * 8 + (r & 7)) * 4 + c) * 2 + (r >> 3)];
        if (c == 0) sp_idx[t * 16 + r] = (uint16_t)(s_packed_meta[t * 8 + (r & 7)] >> (16 * (r >> 3)));
        if (r == 0)
            for (int q = 0; q < 4; ++q) sp_cols[t * 16 + 4 * c + q] = clean_col(s_tile_column[t * 16 + 4 * c + q], ncols);
    }
}

inline unsigned blocks_for(int64_t n, int threads = 256) {
    int64_t b = (n + threads - 1) / threads;
    if (b < 1) b = 1;
    if (b > 65535 * 8) b = 65535 * 8;
    return (unsigned)b;
}

}  // namespace

int Plan::finalize() {
    dev.tc_sub_off = tc_sub_off.as<int32_t>();
    dev.tc_cols = tc_cols.as<int32_t>();
    dev.tc_bits = tc_bits.as<uint64_t>();
    dev.tc_valoff = tc_valoff.as<int32_t>();
    dev.tc_vals = tc_vals.as<uint16_t>();
    dev.sp_tile_off = sp_tile_off.as<int32_t>();
    dev.sp_cols = sp_cols.as<int32_t>();
    dev.sp_a = sp_a.as<uint32_t>();
    dev.sp_idx = sp_idx.as<uint16_t>();
    dev.res_off = res_off.as<int32_t>();
    dev.res_row = res_row.as<int32_t>();
    dev.res_beg = res_beg.as<int32_t>();
    dev.res_cnt = res_cnt.as<int32_t>();
    dev.res_cols = res_cols.as<int32_t>();
    dev.res_vals = res_vals.as<uint16_t>();
    return TRSPMM_OK;
}

int build_plan(const trspmm_meta_view& v, int64_t m_out, int64_t ncols, int dtype, hipStream_t s, Plan& P) {
    TRSPMM_CHECK(v.window == 16, "the routed multiply requires window == 16 (16-row MFMA tiles)");
    TRSPMM_CHECK(dtype == TRSPMM_F16 || dtype == TRSPMM_BF16, "dtype must be TRSPMM_F16 or TRSPMM_BF16");
    TRSPMM_CHECK(m_out > 0 && ncols > 0, "m_out and ncols must be positive");
    TRSPMM_CHECK(v.n_windows >= 0 && v.n_windows * 16 >= m_out, "windows must cover the output rows");
    const int64_t W = v.n_windows;
    P.dev.windows = W;
    P.dev.m_out = m_out;
    P.dev.ncols = ncols;
    P.dev.window = 16;
    P.dev.dtype = dtype;
    P.dev.n_tsub = v.n_tsub;
    P.dev.n_tval = v.n_tval;
    P.dev.n_rparts = v.n_cparts + v.n_csparts;
    P.dev.n_rval = v.n_cval + v.n_csval;

    int rc;
    DevBuf lb_long, lb_short, tile_cnt;
    if ((rc = P.tc_sub_off.alloc(4 * (W + 1))) || (rc = P.res_off.alloc(4 * (W + 1))) ||
        (rc = P.sp_tile_off.alloc(4 * (W + 1))) || (rc = lb_long.alloc(4 * (W + 1))) ||
        (rc = lb_short.alloc(4 * (W + 1))) || (rc = tile_cnt.alloc(4 * (W + 1))))
        return rc;

    hipLaunchKernelGGL(k_window_bounds, dim3(blocks_for(W + 1)), dim3(256), 0, s, v.t_window_row, v.n_tparts,
                       v.t_row_offset, v.c_row, v.n_cparts, v.cs_row, v.n_csparts, W, P.tc_sub_off.as<int32_t>(),
                       lb_long.as<int32_t>(), lb_short.as<int32_t>());
    TRSPMM_HIP(hipGetLastError());

    // TC sub-blocks
    if (v.n_tsub > 0) {
        if ((rc = P.tc_cols.alloc(4 * 8 * v.n_tsub)) || (rc = P.tc_bits.alloc(8 * 2 * v.n_tsub)) ||
            (rc = P.tc_valoff.alloc(4 * (v.n_tsub + 1))) || (rc = P.tc_vals.alloc(2 * (v.n_tval > 0 ? v.n_tval : 1))))
            return rc;
        hipLaunchKernelGGL(k_copy_cols, dim3(blocks_for(8 * v.n_tsub)), dim3(256), 0, s, v.t_column, 8 * v.n_tsub,
                           ncols, P.tc_cols.as<int32_t>());
        TRSPMM_HIP(hipMemcpyAsync(P.tc_bits.p, v.t_binary, 16 * v.n_tsub, hipMemcpyDeviceToDevice, s));
        TRSPMM_HIP(hipMemcpyAsync(P.tc_valoff.p, v.t_block_offset, 4 * (v.n_tsub + 1), hipMemcpyDeviceToDevice, s));
        if (v.n_tval > 0)
            hipLaunchKernelGGL(k_copy_vals, dim3(blocks_for(v.n_tval)), dim3(256), 0, s, v.t_value, v.n_tval,
                               v.t_value_is_f32, dtype, P.tc_vals.as<uint16_t>());
        TRSPMM_HIP(hipGetLastError());
    }

    // residual parts
    const int64_t nparts = v.n_cparts + v.n_csparts;
    if ((rc = P.res_row.alloc(4 * (nparts ? nparts : 1))) || (rc = P.res_beg.alloc(4 * (nparts ? nparts : 1))) ||
        (rc = P.res_cnt.alloc(4 * (nparts ? nparts : 1))))
        return rc;
    if (W > 0) {
        hipLaunchKernelGGL(k_residual_parts, dim3((unsigned)W), dim3(64), 0, s, lb_long.as<int32_t>(),
                           lb_short.as<int32_t>(), v.c_row_offset, v.c_row, v.cs_row_offset, v.cs_row, v.n_cval, W,
                           P.res_off.as<int32_t>(), P.res_row.as<int32_t>(), P.res_beg.as<int32_t>(),
                           P.res_cnt.as<int32_t>());
        TRSPMM_HIP(hipGetLastError());
    } else {
        TRSPMM_HIP(hipMemsetAsync(P.res_off.p, 0, 4, s));
    }
    const int64_t nres = v.n_cval + v.n_csval;
    if (nres > 0) {
        if ((rc = P.res_cols.alloc(4 * nres)) || (rc = P.res_vals.alloc(2 * nres))) return rc;
        if (v.n_cval > 0) {
            hipLaunchKernelGGL(k_copy_cols, dim3(blocks_for(v.n_cval)), dim3(256), 0, s, v.c_column, v.n_cval, ncols,
                               P.res_cols.as<int32_t>());
            hipLaunchKernelGGL(k_copy_vals, dim3(blocks_for(v.n_cval)), dim3(256), 0, s, v.c_value, v.n_cval,
                               v.c_value_is_f32, dtype, P.res_vals.as<uint16_t>());
        }
        if (v.n_csval > 0) {
            hipLaunchKernelGGL(k_copy_cols, dim3(blocks_for(v.n_csval)), dim3(256), 0, s, v.cs_column, v.n_csval,
                               ncols, P.res_cols.as<int32_t>() + v.n_cval);
            hipLaunchKernelGGL(k_copy_vals, dim3(blocks_for(v.n_csval)), dim3(256), 0, s, v.cs_value, v.n_csval,
                               v.c_value_is_f32, dtype, P.res_vals.as<uint16_t>() + v.n_cval);
        }
        TRSPMM_HIP(hipGetLastError());
    }

    // 2:4 tiles
    if (W > 0) {
        hipLaunchKernelGGL(k_tile_counts, dim3(blocks_for(W)), dim3(256), 0, s, v.s_offsets, W, tile_cnt.as<int32_t>());
        TRSPMM_HIP(hipGetLastError());
    }
    if ((rc = exclusive_scan<int32_t>(W > 0 ? tile_cnt.as<int32_t>() : nullptr, P.sp_tile_off.as<int32_t>(), W, s)))
        return rc;
    int32_t n_tiles = 0;
    TRSPMM_HIP(hipMemcpyAsync(&n_tiles, P.sp_tile_off.as<int32_t>() + W, 4, hipMemcpyDeviceToHost, s));
    TRSPMM_HIP(hipStreamSynchronize(s));
    P.dev.n_tiles = n_tiles;
    if (n_tiles > 0) {
        if ((rc = P.sp_cols.alloc(4 * 16 * (int64_t)n_tiles)) || (rc = P.sp_a.alloc(4 * 64 * (int64_t)n_tiles)) ||
            (rc = P.sp_idx.alloc(2 * 16 * (int64_t)n_tiles)))
            return rc;
        hipLaunchKernelGGL(k_tiles_from_groups, dim3((unsigned)W), dim3(256), 0, s, v.s_offsets, v.s_column,
                           v.s_value, v.s_pos, P.sp_tile_off.as<int32_t>(), W, v.window, ncols, dtype,
                           P.sp_cols.as<int32_t>(), P.sp_a.as<uint32_t>(), P.sp_idx.as<uint16_t>());
        TRSPMM_HIP(hipGetLastError());
    }
    TRSPMM_HIP(hipStreamSynchronize(s));   // temporaries are freed on return
    return P.finalize();
}

int build_plan_packed(const int32_t* s_tile_column, const int32_t* s_packed_a, const int32_t* s_packed_meta,
                      const int32_t* s_window_tile_offset, int64_t T, int64_t W, int window, int64_t m_out,
                      int64_t ncols, hipStream_t s, Plan& P) {
    TRSPMM_CHECK(window == 16, "the routed multiply requires window == 16 (16-row MFMA tiles)");
    TRSPMM_CHECK(m_out > 0 && ncols > 0, "m_out and ncols must be positive");
    TRSPMM_CHECK(W * 16 >= m_out, "windows must cover the output rows");
    P.dev.windows = W;
    P.dev.m_out = m_out;
    P.dev.ncols = ncols;
    P.dev.dtype = TRSPMM_F16;
    P.dev.n_tiles = T;
    int rc;
    if ((rc = P.tc_sub_off.alloc(4 * (W + 1))) || (rc = P.res_off.alloc(4 * (W + 1))) ||
        (rc = P.sp_tile_off.alloc(4 * (W + 1))))
        return rc;
    TRSPMM_HIP(hipMemsetAsync(P.tc_sub_off.p, 0, 4 * (W + 1), s));
    TRSPMM_HIP(hipMemsetAsync(P.res_off.p, 0, 4 * (W + 1), s));
    TRSPMM_HIP(hipMemcpyAsync(P.sp_tile_off.p, s_window_tile_offset, 4 * (W + 1), hipMemcpyDeviceToDevice, s));
    if (T > 0) {
        if ((rc = P.sp_cols.alloc(4 * 16 * T)) || (rc = P.sp_a.alloc(4 * 64 * T)) || (rc = P.sp_idx.alloc(2 * 16 * T)))
            return rc;
        hipLaunchKernelGGL(k_tiles_from_packed, dim3(blocks_for(T * 64)), dim3(256), 0, s, s_tile_column,
                           (const uint32_t*)s_packed_a, (const uint32_t*)s_packed_meta, T, ncols,
                           P.sp_cols.as<int32_t>(), P.sp_a.as<uint32_t>(), P.sp_idx.as<uint16_t>());
        TRSPMM_HIP(hipGetLastError());
    }
    TRSPMM_HIP(hipStreamSynchronize(s));
    return P.finalize();
}

}  // namespace trspmm
