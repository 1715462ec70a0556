// The legacy Libra kernel family (TR-source/SpMM/mGCNkernel.cu:1357-3837):
// host functions that multiply with the original Libra metadata, fp16 and tf32.
//
// Format, as the reference's kernels read it:
//   * TC route (spmm_forward_fp16_csr_v2_kernel_tcu, mGCNkernel.cu:15-212):
//     8-row windows; part p covers blocks t_row_offset[p] .. [p + 1] of
//     window t_window_row[p] (output rows 8 t_window_row[p] + n, :132); block
//     b is 8 columns t_column[8 b + k] (-1 empty) and a 64-bit mask
//     t_binary[b] whose bit 8 n + k marks row n, column k (lane l reads bits
//     2 l, 2 l + 1 as the m16n8k8 B-fragment rows (l % 4) 2 + j of column
//     l / 4, :73-95); its values sit at t_blockNew_offset[b] in bit order;
//   * residual route (spmm_forward_fp16_csr_v2_kernel_cuda{,_short},
//     :977-1043, :1120-1171): part p adds c_value[e] B[c_column[e]] over
//     entries c_row_offset[p] .. [p + 1] to row c_row[p], the same layout as
//     the online router's residual lists.
// t_atomic / c_atomic say whether a part stores or adds; every row here has
// one owner, so they are read by nobody.
//
// The tf32 members (spmm_forward_tf32_csr_v2_kernel_tcu, :2886-3030) use
// 8 x 4 blocks: t_column[4 b + k], a 32-bit mask whose bit 4 n + k (lane l
// of the m16n8k4 B fragment, row k = l % 4 of column n = l / 4) is row n,
// column k; values, residual values and B are fp32.  gfx950 has no tf32
// (xf32) matrix instruction, so these run as three bf16 products,
// C = A_hi B_hi + A_hi B_lo + A_lo B_hi with x_hi = bf16(x) and
// x_lo = bf16(x - x_hi) (the A_lo B_lo term is below 2^-16 of the product):
// more precise than tf32's 10-bit mantissa.
//
// The plan is built from the same plan builder as the online ops: an 8-row
// block becomes the upper or lower half of a 16-row TC sub-block (tcu
// variants: matrix cores for TC, the tile kernel's residual engine for the
// rest), or every nonzero becomes a residual record of its row (cc_only: the
// vector-ALU residual kernel for everything, the reference's CUDA-core
// emulation of the TC route, :176-212).  Each entry point builds the plan,
// runs the reference's 10 untimed launches and `epoches` timed ones on the
// null stream, leaves the product of the last launch in output_matrix and
// returns the mean ms per launch (-1 on error, message in trspmm_last_error).
#include <algorithm>
#include <vector>

#include "plan.h"
#include "scan.h"

namespace trspmm {
namespace {

// legacy TC -> the plan builder's TC layout: a 16-row window per two 8-row
// windows (t_window_row = 16 * (w8 / 2), the block's mask in half w8 % 2), or
// for the all-residual build an 8-row window (t_window_row = 8 w8, low half)
__global__ void k_libra_tc(const int32_t* __restrict__ t_row_offset, const int32_t* __restrict__ t_window_row,
                           const int64_t* __restrict__ t_binary, int64_t parts, int64_t blocks, int window,
                           int32_t* __restrict__ wrow, int64_t* __restrict__ bits2) {
    const int64_t n = parts > blocks ? parts : blocks;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
    if (i < parts) {
        const int64_t w8 = t_window_row[i];
        wrow[i] = window == 16 ? (int32_t)(16 * (w8 / 2)) : (int32_t)(8 * w8);
    }
    if (i < blocks) {
        // the block's part: last p with t_row_offset[p] <= i
        int64_t lo = 0, hi = parts;
        while (lo < hi) {
            const int64_t mid = (lo + hi + 1) / 2;
            if (t_row_offset[mid] <= i) lo = mid;
            else hi = mid - 1;
        }
        const int half = window == 16 ? (int)(t_window_row[lo] & 1) : 0;
        bits2[2 * i + half] = t_binary[i];
        bits2[2 * i + 1 - half] = 0;
    }
    }
}

// tf32 blocks (8 x 4, 32-bit mask bit 4 n + k) -> 8 x 8 blocks (bit 8 n + k, the
// columns k >= 4 empty), then placed as k_libra_tc places 8 x 8 blocks
__global__ void k_libra_tc32(const int32_t* __restrict__ t_row_offset, const int32_t* __restrict__ t_window_row,
                             const int32_t* __restrict__ t_binary32, const int32_t* __restrict__ t_column4,
                             int64_t parts, int64_t blocks, int window, int32_t* __restrict__ wrow,
                             int64_t* __restrict__ bits2, int32_t* __restrict__ cols8) {
    const int64_t n = parts > blocks ? parts : blocks;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
    if (i < parts) {
        const int64_t w8 = t_window_row[i];
        wrow[i] = window == 16 ? (int32_t)(16 * (w8 / 2)) : (int32_t)(8 * w8);
    }
    if (i < blocks) {
        int64_t lo = 0, hi = parts;
        while (lo < hi) {
            const int64_t mid = (lo + hi + 1) / 2;
            if (t_row_offset[mid] <= i) lo = mid;
            else hi = mid - 1;
        }
        const uint32_t m32 = (uint32_t)t_binary32[i];
        uint64_t m64 = 0;
        for (int l = 0; l < 32; ++l)
            if ((m32 >> l) & 1u) m64 |= 1ull << (8 * (l / 4) + (l % 4));
        const int half = window == 16 ? (int)(t_window_row[lo] & 1) : 0;
        bits2[2 * i + half] = (int64_t)m64;
        bits2[2 * i + 1 - half] = 0;
        for (int k = 0; k < 8; ++k) cols8[8 * i + k] = k < 4 ? t_column4[4 * i + k] : -1;
    }
    }
}

// x -> (x, x - bf16(x)): the low part of the three-term bf16 split
__global__ void k_lo_part(const float* __restrict__ x, int64_t n, float* __restrict__ lo) {
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
        const float v = x[i];
        lo[i] = v - __builtin_bit_cast(float, (uint32_t)f32_to_bf16_bits(v) << 16);
    }
}

// B fp32 [rows, dimN] -> B_hi, B_lo bf16 [rows, K8] (zero padded)
__global__ void k_split_rhs(const float* __restrict__ B, int64_t rows, int dimN, int64_t K8, uint16_t* __restrict__ hi,
                            uint16_t* __restrict__ lo) {
    const int64_t n = rows * K8;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
        const int64_t r = i / K8, f = i % K8;
        const float v = f < dimN ? B[r * dimN + f] : 0.f;
        const uint16_t h = f32_to_bf16_bits(v);
        hi[i] = h;
        lo[i] = f32_to_bf16_bits(v - __builtin_bit_cast(float, (uint32_t)h << 16));
    }
}

// C[i] += T1[i] + T2[i]
__global__ void k_add2(float* __restrict__ C, const float* __restrict__ T1, const float* __restrict__ T2, int64_t n) {
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
        C[i] += T1[i] + T2[i];
}

// largest column index of a list (for the operand's row count), into *out
__global__ void k_max_col(const int32_t* __restrict__ col, int64_t n, int32_t* __restrict__ out) {
    int32_t m = -1;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
        m = col[i] > m ? col[i] : m;
    for (int off = 32; off > 0; off >>= 1) {
        const int32_t o = __shfl_xor(m, off);
        m = o > m ? o : m;
    }
    if ((threadIdx.x & 63) == 0) atomicMax(out, m);
}

unsigned grid_for(int64_t n) {
    int64_t b = (n + 255) / 256;
    return (unsigned)std::max<int64_t>(1, std::min<int64_t>(b, 65536));
}

template <typename T>
int to_host(std::vector<T>& h, const T* d, int64_t n) {
    h.resize((size_t)std::max<int64_t>(n, 0));
    if (n > 0) TRSPMM_HIP(hipMemcpy(h.data(), d, sizeof(T) * (size_t)n, hipMemcpyDeviceToHost));
    return TRSPMM_OK;
}

int check_offsets(const std::vector<int32_t>& off, const char* name) {
    for (size_t i = 1; i < off.size(); ++i)
        TRSPMM_CHECK(off[i] >= off[i - 1], std::string(name) + " must be non-decreasing");
    TRSPMM_CHECK(off.empty() || off[0] == 0, std::string(name) + " must start at 0");
    return TRSPMM_OK;
}

// fp16 members: values and B 16-bit, masks int64 over 8 columns; tf32 members:
// values and B fp32, masks int32 over 4 columns
struct LibraTc {
    const int32_t *t_row_offset, *t_blockNew_offset, *t_column, *t_window_row;
    const void* t_value;
    const void* t_binary;
    int64_t parts;
};
struct LibraRes {
    const int32_t *row_offset, *row, *column;
    const void* value;
    int64_t parts;
};

// engine: 0 = TC on matrix cores, the residual lists in the tile kernel;
// 1 = every nonzero a residual record on the vector-ALU kernel
int libra_run(const LibraTc* tc, const LibraRes* res, const LibraRes* res_short, const void* rhs,
              float* out, int dimN, int mOri, int epoches, int engine, float* ms, bool tf32 = false) {
    TRSPMM_CHECK(dimN > 0 && mOri > 0 && epoches > 0, "dimN, mOri and epoches must be positive");
    TRSPMM_CHECK(rhs && out, "rhs_matrix and output_matrix must not be null");
    const hipStream_t s = nullptr;
    StreamArena A(s);
    int rc;
    // sizes from the offsets (device arrays), checked on the host
    std::vector<int32_t> t_off, t_wrow, t_voff, c_off, c_row, cs_off, cs_row;
    int64_t blocks = 0, n_tval = 0, n_c = 0, n_cs = 0, rows_needed = mOri;
    if (tc && tc->parts > 0) {
        TRSPMM_CHECK(tc->t_row_offset && tc->t_window_row && tc->t_blockNew_offset && tc->t_column && tc->t_binary &&
                         tc->t_value, "TC metadata pointers must not be null");
        if ((rc = to_host(t_off, tc->t_row_offset, tc->parts + 1)) || (rc = check_offsets(t_off, "t_row_offset")) ||
            (rc = to_host(t_wrow, tc->t_window_row, tc->parts)))
            return rc;
        blocks = t_off.back();
        for (size_t p = 0; p < t_wrow.size(); ++p) {
            TRSPMM_CHECK(t_wrow[p] >= 0 && (p == 0 || t_wrow[p] >= t_wrow[p - 1]),
                         "t_window_row must be non-negative and non-decreasing");
            rows_needed = std::max<int64_t>(rows_needed, 8 * (int64_t)t_wrow[p] + 8);
        }
        if ((rc = to_host(t_voff, tc->t_blockNew_offset, blocks + 1)) || (rc = check_offsets(t_voff, "t_blockNew_offset")))
            return rc;
        n_tval = t_voff.back();
    }
    const LibraRes* lists[2] = {res, res_short};
    std::vector<int32_t>* offs[2] = {&c_off, &cs_off};
    std::vector<int32_t>* rws[2] = {&c_row, &cs_row};
    int64_t* counts[2] = {&n_c, &n_cs};
    for (int k = 0; k < 2; ++k) {
        const LibraRes* L = lists[k];
        if (!L || L->parts <= 0) continue;
        TRSPMM_CHECK(L->row_offset && L->row && L->column && L->value, "residual metadata pointers must not be null");
        if ((rc = to_host(*offs[k], L->row_offset, L->parts + 1)) ||
            (rc = check_offsets(*offs[k], k ? "c_row_offset_short" : "c_row_offset")) ||
            (rc = to_host(*rws[k], L->row, L->parts)))
            return rc;
        *counts[k] = offs[k]->back();
        for (int32_t r : *rws[k]) {
            TRSPMM_CHECK(r >= 0, "c_row entries must be non-negative");
            rows_needed = std::max<int64_t>(rows_needed, (int64_t)r + 1);
        }
    }
    // rows of B: one past the largest column the metadata names
    DevBuf mx;
    if ((rc = mx.alloc(A, 4))) return rc;
    TRSPMM_HIP(hipMemsetAsync(mx.p, 0xff, 4, s));
    const int tcw = tf32 ? 4 : 8;                      // columns per block
    if (blocks > 0)
        hipLaunchKernelGGL(k_max_col, dim3(grid_for(tcw * blocks)), dim3(256), 0, s, tc->t_column, tcw * blocks,
                           mx.as<int32_t>());
    if (n_c > 0) hipLaunchKernelGGL(k_max_col, dim3(grid_for(n_c)), dim3(256), 0, s, res->column, n_c, mx.as<int32_t>());
    if (n_cs > 0)
        hipLaunchKernelGGL(k_max_col, dim3(grid_for(n_cs)), dim3(256), 0, s, res_short->column, n_cs, mx.as<int32_t>());
    TRSPMM_HIP(hipGetLastError());
    int32_t maxcol = -1;
    TRSPMM_HIP(hipMemcpyAsync(&maxcol, mx.p, 4, hipMemcpyDeviceToHost, s));
    TRSPMM_HIP(hipStreamSynchronize(s));
    const int64_t ncols = std::max<int32_t>(maxcol, 0) + 1;

    const int window = engine == 1 ? 8 : 16;
    const int64_t W = (rows_needed + window - 1) / window;
    DevBuf wrow, bits2, cols8, zeros;
    if ((rc = zeros.alloc(A, 4 * (W + 2)))) return rc;
    TRSPMM_HIP(hipMemsetAsync(zeros.p, 0, 4 * (W + 2), s));
    trspmm_meta_view v{};
    v.window = window;
    v.n_windows = W;
    v.t_value_is_f32 = tf32 ? 1 : 0;
    v.c_value_is_f32 = tf32 ? 1 : 0;
    v.t_row_offset = v.c_row_offset = v.cs_row_offset = zeros.as<int32_t>();
    v.s_offsets = zeros.as<int32_t>();
    if (blocks > 0) {
        if ((rc = wrow.alloc(A, 4 * tc->parts)) || (rc = bits2.alloc(A, 16 * blocks))) return rc;
        if (tf32) {
            if ((rc = cols8.alloc(A, 32 * blocks))) return rc;
            hipLaunchKernelGGL(k_libra_tc32, dim3(grid_for(std::max<int64_t>(tc->parts, blocks))), dim3(256), 0, s,
                               tc->t_row_offset, tc->t_window_row, (const int32_t*)tc->t_binary, tc->t_column,
                               tc->parts, blocks, window, wrow.as<int32_t>(), bits2.as<int64_t>(), cols8.as<int32_t>());
        } else {
            hipLaunchKernelGGL(k_libra_tc, dim3(grid_for(std::max<int64_t>(tc->parts, blocks))), dim3(256), 0, s,
                               tc->t_row_offset, tc->t_window_row, (const int64_t*)tc->t_binary, tc->parts, blocks,
                               window, wrow.as<int32_t>(), bits2.as<int64_t>());
        }
        TRSPMM_HIP(hipGetLastError());
        v.t_row_offset = tc->t_row_offset;
        v.n_tparts = tc->parts;
        v.t_block_offset = tc->t_blockNew_offset;
        v.t_column = tf32 ? cols8.as<int32_t>() : tc->t_column;
        v.n_tsub = blocks;
        v.t_value = tc->t_value;
        v.n_tval = n_tval;
        v.t_window_row = wrow.as<int32_t>();
        v.t_binary = bits2.as<int64_t>();
    }
    if (n_c > 0 || (res && res->parts > 0)) {
        v.c_row_offset = res->row_offset;
        v.n_cparts = res->parts;
        v.c_row = res->row;
        v.c_column = res->column;
        v.n_cval = n_c;
        v.c_value = res->value;
    }
    if (n_cs > 0 || (res_short && res_short->parts > 0)) {
        v.cs_row_offset = res_short->row_offset;
        v.n_csparts = res_short->parts;
        v.cs_row = res_short->row;
        v.cs_column = res_short->column;
        v.n_csval = n_cs;
        v.cs_value = res_short->value;
    }
    Plan P;
    P.opts.residual_engine = engine == 1 ? TRSPMM_ENGINE_VALU : TRSPMM_ENGINE_MFMA;
    P.opts.k_hint = dimN;
    if ((rc = build_plan(v, mOri, ncols, tf32 ? TRSPMM_BF16 : TRSPMM_F16, s, P))) return rc;

    // operands of 8-feature rows (the kernels read 16-byte pieces): B and C are
    // used in place when dimN is a multiple of 8, else through padded copies
    const int64_t K8 = (dimN + 7) / 8 * 8;
    const uint16_t* B = (const uint16_t*)rhs;
    float* C = out;
    DevBuf Bp, Cp, Blo, T1, T2, vlo[3];
    Plan PL;                                           // tf32: the plan of the values' low parts
    if (tf32) {
        // B_hi, B_lo and the values' low parts; the low plan is the same routing
        if ((rc = Bp.alloc(A, 2 * ncols * K8)) || (rc = Blo.alloc(A, 2 * ncols * K8)) ||
            (rc = T1.alloc(A, 4 * (int64_t)mOri * K8)) || (rc = T2.alloc(A, 4 * (int64_t)mOri * K8)))
            return rc;
        hipLaunchKernelGGL(k_split_rhs, dim3(grid_for(ncols * K8)), dim3(256), 0, s, (const float*)rhs, ncols, dimN, K8,
                           Bp.as<uint16_t>(), Blo.as<uint16_t>());
        TRSPMM_HIP(hipGetLastError());
        trspmm_meta_view vl = v;
        const void* src[3] = {v.t_value, v.c_value, v.cs_value};
        const int64_t cnt[3] = {v.n_tval, v.n_cval, v.n_csval};
        const void** dst[3] = {&vl.t_value, &vl.c_value, &vl.cs_value};
        for (int k = 0; k < 3; ++k) {
            if (cnt[k] <= 0) continue;
            if ((rc = vlo[k].alloc(A, 4 * cnt[k]))) return rc;
            hipLaunchKernelGGL(k_lo_part, dim3(grid_for(cnt[k])), dim3(256), 0, s, (const float*)src[k], cnt[k],
                               vlo[k].as<float>());
            TRSPMM_HIP(hipGetLastError());
            *dst[k] = vlo[k].p;
        }
        PL.opts = P.opts;
        if ((rc = build_plan(vl, mOri, ncols, TRSPMM_BF16, s, PL))) return rc;
        B = Bp.as<uint16_t>();
        if ((uintptr_t)out & 15 || K8 != dimN) {
            if ((rc = Cp.alloc(A, 4 * (int64_t)mOri * K8))) return rc;
            C = Cp.as<float>();
        }
    } else if (K8 != dimN || ((uintptr_t)rhs & 15) || ((uintptr_t)out & 15)) {
        if ((rc = Bp.alloc(A, 2 * ncols * K8)) || (rc = Cp.alloc(A, 4 * (int64_t)mOri * K8))) return rc;
        TRSPMM_HIP(hipMemsetAsync(Bp.p, 0, 2 * ncols * K8, s));
        TRSPMM_HIP(hipMemcpy2DAsync(Bp.p, 2 * K8, rhs, 2 * (size_t)dimN, 2 * (size_t)dimN, (size_t)ncols,
                                    hipMemcpyDeviceToDevice, s));
        B = Bp.as<uint16_t>();
        C = Cp.as<float>();
    }
    // one multiply: the plan's product, or for tf32 the three bf16 products summed
    auto multiply = [&]() -> int {
        int r = launch_spmm(P.dev, B, K8, K8, C, K8, TRSPMM_ROUTE_ALL, s);
        if (r || !tf32) return r;
        if ((r = launch_spmm(P.dev, Blo.as<uint16_t>(), K8, K8, T1.as<float>(), K8, TRSPMM_ROUTE_ALL, s)) ||
            (r = launch_spmm(PL.dev, B, K8, K8, T2.as<float>(), K8, TRSPMM_ROUTE_ALL, s)))
            return r;
        hipLaunchKernelGGL(k_add2, dim3(grid_for((int64_t)mOri * K8)), dim3(256), 0, s, C, T1.as<float>(),
                           T2.as<float>(), (int64_t)mOri * K8);
        TRSPMM_HIP(hipGetLastError());
        return TRSPMM_OK;
    };
    for (int i = 0; i < 10; ++i)
        if ((rc = multiply())) return rc;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    TRSPMM_HIP(hipEventCreate(&e0));
    if (hipEventCreate(&e1) != hipSuccess) {
        (void)hipEventDestroy(e0);
        set_error("hipEventCreate failed");
        return TRSPMM_EHIP;
    }
    rc = TRSPMM_OK;
    if (hipEventRecord(e0, s) != hipSuccess) rc = TRSPMM_EHIP;
    for (int i = 0; i < epoches && rc == TRSPMM_OK; ++i) rc = multiply();
    if (rc == TRSPMM_OK && hipEventRecord(e1, s) != hipSuccess) rc = TRSPMM_EHIP;
    float total = 0.f;
    if (rc == TRSPMM_OK && (hipEventSynchronize(e1) != hipSuccess || hipEventElapsedTime(&total, e0, e1) != hipSuccess))
        rc = TRSPMM_EHIP;
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    if (rc != TRSPMM_OK) {
        if (rc == TRSPMM_EHIP) set_error("timing events failed");
        return rc;
    }
    if (C != out)
        TRSPMM_HIP(hipMemcpy2DAsync(out, 4 * (size_t)dimN, C, 4 * K8, 4 * (size_t)dimN, (size_t)mOri,
                                    hipMemcpyDeviceToDevice, s));
    TRSPMM_HIP(hipStreamSynchronize(s));
    *ms = total / (float)epoches;
    return TRSPMM_OK;
}

float libra_entry(const LibraTc* tc, const LibraRes* res, const LibraRes* res_short, const void* rhs, float* out,
                  int dimN, int mOri, int epoches, int engine, bool tf32 = false) {
    float ms = -1.f;
    const int rc = libra_run(tc, res, res_short, rhs, out, dimN, mOri, epoches, engine, &ms, tf32);
    return rc == TRSPMM_OK ? ms : -1.f;
}

}  // namespace
}  // namespace trspmm

using trspmm::LibraRes;
using trspmm::LibraTc;

#define TRSPMM_LIBRA_FULL_ARGS                                                                                      \
    const int32_t *t_row_offset, const int32_t *t_blockNew_offset, const uint16_t *t_value, const int32_t *t_column, \
        const int32_t *t_window_row, const int32_t *t_atomic, const int64_t *t_binary, const int32_t *c_row_offset,   \
        const int32_t *c_row, const int32_t *c_atomic, const int32_t *c_column, const uint16_t *c_value,              \
        const int32_t *c_row_offset_short, const int32_t *c_row_short, const int32_t *c_atomic_short,                 \
        const int32_t *c_column_short, const uint16_t *c_value_short, const uint16_t *rhs_matrix,                     \
        float *output_matrix, int parts_t, int parts_c, int partsize_c, int parts_c_short, int dimN, int mOri,       \
        int epoches

static float libra_full(TRSPMM_LIBRA_FULL_ARGS, int engine) {
    (void)t_atomic, (void)c_atomic, (void)c_atomic_short, (void)partsize_c;
    const LibraTc tc{t_row_offset, t_blockNew_offset, t_column, t_window_row, t_value, t_binary, parts_t};
    const LibraRes r{c_row_offset, c_row, c_column, c_value, parts_c};
    const LibraRes rs{c_row_offset_short, c_row_short, c_column_short, c_value_short, parts_c_short};
    return trspmm::libra_entry(&tc, &r, &rs, rhs_matrix, output_matrix, dimN, mOri, epoches, engine);
}

extern "C" {

float trspmm_libra_fp16_tcu_cuda(TRSPMM_LIBRA_FULL_ARGS) {
    return libra_full(t_row_offset, t_blockNew_offset, t_value, t_column, t_window_row, t_atomic, t_binary,
                      c_row_offset, c_row, c_atomic, c_column, c_value, c_row_offset_short, c_row_short,
                      c_atomic_short, c_column_short, c_value_short, rhs_matrix, output_matrix, parts_t, parts_c,
                      partsize_c, parts_c_short, dimN, mOri, epoches, 0);
}

float trspmm_libra_fp16_tcu_cuda_seq(TRSPMM_LIBRA_FULL_ARGS) {
    return libra_full(t_row_offset, t_blockNew_offset, t_value, t_column, t_window_row, t_atomic, t_binary,
                      c_row_offset, c_row, c_atomic, c_column, c_value, c_row_offset_short, c_row_short,
                      c_atomic_short, c_column_short, c_value_short, rhs_matrix, output_matrix, parts_t, parts_c,
                      partsize_c, parts_c_short, dimN, mOri, epoches, 0);
}

float trspmm_libra_fp16_tcu_cuda_hybrid(TRSPMM_LIBRA_FULL_ARGS) {
    return libra_full(t_row_offset, t_blockNew_offset, t_value, t_column, t_window_row, t_atomic, t_binary,
                      c_row_offset, c_row, c_atomic, c_column, c_value, c_row_offset_short, c_row_short,
                      c_atomic_short, c_column_short, c_value_short, rhs_matrix, output_matrix, parts_t, parts_c,
                      partsize_c, parts_c_short, dimN, mOri, epoches, 0);
}

float trspmm_libra_fp16_tcu_cuda_hybrid_seq(TRSPMM_LIBRA_FULL_ARGS) {
    return libra_full(t_row_offset, t_blockNew_offset, t_value, t_column, t_window_row, t_atomic, t_binary,
                      c_row_offset, c_row, c_atomic, c_column, c_value, c_row_offset_short, c_row_short,
                      c_atomic_short, c_column_short, c_value_short, rhs_matrix, output_matrix, parts_t, parts_c,
                      partsize_c, parts_c_short, dimN, mOri, epoches, 0);
}

float trspmm_libra_fp16_tcu_cuda_cc_only(TRSPMM_LIBRA_FULL_ARGS) {
    return libra_full(t_row_offset, t_blockNew_offset, t_value, t_column, t_window_row, t_atomic, t_binary,
                      c_row_offset, c_row, c_atomic, c_column, c_value, c_row_offset_short, c_row_short,
                      c_atomic_short, c_column_short, c_value_short, rhs_matrix, output_matrix, parts_t, parts_c,
                      partsize_c, parts_c_short, dimN, mOri, epoches, 1);
}

float trspmm_libra_fp16_tcu_nomap(const int32_t* t_row_offset, const int32_t* t_blockNew_offset,
                                  const uint16_t* t_value, const int32_t* t_column, const int32_t* t_window_row,
                                  const int32_t* t_atomic, const int64_t* t_binary, const uint16_t* rhs_matrix,
                                  float* output_matrix, int parts_t, int dimN, int mOri, int epoches) {
    (void)t_atomic;
    const LibraTc tc{t_row_offset, t_blockNew_offset, t_column, t_window_row, t_value, t_binary, parts_t};
    return trspmm::libra_entry(&tc, nullptr, nullptr, rhs_matrix, output_matrix, dimN, mOri, epoches, 0);
}

float trspmm_libra_fp16_cuda_split(const int32_t* c_row_offset, const int32_t* c_row, const int32_t* c_atomic,
                                   const int32_t* c_column, const uint16_t* c_value, const int32_t* c_row_offset_short,
                                   const int32_t* c_row_short, const int32_t* c_atomic_short,
                                   const int32_t* c_column_short, const uint16_t* c_value_short,
                                   const uint16_t* rhs_matrix, float* output_matrix, int partsize, int dimN, int mOri,
                                   int epoches, int parts, int parts_short, int swizzle) {
    (void)c_atomic, (void)c_atomic_short, (void)partsize, (void)swizzle;
    const LibraRes r{c_row_offset, c_row, c_column, c_value, parts};
    const LibraRes rs{c_row_offset_short, c_row_short, c_column_short, c_value_short, parts_short};
    return trspmm::libra_entry(nullptr, &r, &rs, rhs_matrix, output_matrix, dimN, mOri, epoches, 0);
}

float trspmm_libra_fp16_cuda_split_v1(const int32_t* c_row_offset, const int32_t* c_row, const int32_t* c_atomic,
                                      const int32_t* c_column, const uint16_t* c_value,
                                      const int32_t* c_row_offset_short, const int32_t* c_row_short,
                                      const int32_t* c_atomic_short, const int32_t* c_column_short,
                                      const uint16_t* c_value_short, const uint16_t* rhs_matrix, float* output_matrix,
                                      int partsize, int dimN, int mOri, int epoches, int parts, int parts_short,
                                      int swizzle) {
    return trspmm_libra_fp16_cuda_split(c_row_offset, c_row, c_atomic, c_column, c_value, c_row_offset_short,
                                        c_row_short, c_atomic_short, c_column_short, c_value_short, rhs_matrix,
                                        output_matrix, partsize, dimN, mOri, epoches, parts, parts_short, swizzle);
}

#define TRSPMM_LIBRA_TF32_ARGS                                                                                    \
    const int32_t *t_row_offset, const int32_t *t_blockNew_offset, const float *t_value, const int32_t *t_column,    \
        const int32_t *t_window_row, const int32_t *t_atomic, const int32_t *t_binary, const int32_t *c_row_offset,   \
        const int32_t *c_row, const int32_t *c_atomic, const int32_t *c_column, const float *c_value,                 \
        const int32_t *c_row_offset_short, const int32_t *c_row_short, const int32_t *c_atomic_short,                 \
        const int32_t *c_column_short, const float *c_value_short, const float *rhs_matrix, float *output_matrix,     \
        int parts_t, int parts_c, int partsize_c, int parts_c_short, int dimN, int mOri, int epoches

float trspmm_libra_tf32_tcu_cuda(TRSPMM_LIBRA_TF32_ARGS) {
    (void)t_atomic, (void)c_atomic, (void)c_atomic_short, (void)partsize_c;
    const LibraTc tc{t_row_offset, t_blockNew_offset, t_column, t_window_row, t_value, t_binary, parts_t};
    const LibraRes r{c_row_offset, c_row, c_column, c_value, parts_c};
    const LibraRes rs{c_row_offset_short, c_row_short, c_column_short, c_value_short, parts_c_short};
    return trspmm::libra_entry(&tc, &r, &rs, rhs_matrix, output_matrix, dimN, mOri, epoches, 0, true);
}

float trspmm_libra_tf32_tcu_cuda_seq(TRSPMM_LIBRA_TF32_ARGS) {
    return trspmm_libra_tf32_tcu_cuda(t_row_offset, t_blockNew_offset, t_value, t_column, t_window_row, t_atomic,
                                      t_binary, c_row_offset, c_row, c_atomic, c_column, c_value, c_row_offset_short,
                                      c_row_short, c_atomic_short, c_column_short, c_value_short, rhs_matrix,
                                      output_matrix, parts_t, parts_c, partsize_c, parts_c_short, dimN, mOri, epoches);
}

float trspmm_libra_tf32_cuda_split(const int32_t* c_row_offset, const int32_t* c_row, const int32_t* c_atomic,
                                   const int32_t* c_column, const float* c_value, const int32_t* c_row_offset_short,
                                   const int32_t* c_row_short, const int32_t* c_atomic_short,
                                   const int32_t* c_column_short, const float* c_value_short, const float* rhs_matrix,
                                   float* output_matrix, int partsize, int dimN, int mOri, int epoches, int parts,
                                   int parts_short, int swizzle) {
    (void)c_atomic, (void)c_atomic_short, (void)partsize, (void)swizzle;
    const LibraRes r{c_row_offset, c_row, c_column, c_value, parts};
    const LibraRes rs{c_row_offset_short, c_row_short, c_column_short, c_value_short, parts_short};
    return trspmm::libra_entry(nullptr, &r, &rs, rhs_matrix, output_matrix, dimN, mOri, epoches, 0, true);
}

}  // extern "C"
