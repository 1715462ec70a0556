// Device plan builder: reference-format routed metadata -> gfx950 lane layouts.
//
// Input contract (the op's inputs, TR-source/SpMM/mGCN_online.cpp:760-932):
// TC parts carry t_window_row = 16*window (block_online.cpp:493) and are in
// window order; residual parts (long and short lists) are in row order; 2:4
// groups are in window order with s_offsets[W+1].  The plan keeps the same
// nonzeros and values; only layouts change:
//   * per-window ranges of TC sub-blocks and 2:4 tiles,
//   * 2:4 tiles repacked so a lane reads its row's 4 groups as 16 contiguous
//     bytes plus one 16-bit index word (the SMFMAC 16x16x64 A-operand form),
//   * the residual parts (long list + short list) merged into one CSR over
//     rows: a row's long-part entries, then its short-part entries, which is
//     the row's CSR order (block_online.cpp:415-439),
//   * values converted once to the plan's 16-bit type (fp16 or bf16),
//   * columns outside [0, ncols) replaced by -1 (contribute nothing), the
//     bounds rule the reference kernels apply (mGCNkernel_online.cu:296-306),
//   * work-item lists for the tile-route and residual-route kernels, the
//     tile items ordered heaviest first (except for item_order 1 and 3).
#include <algorithm>
#include <cstdlib>
#include <cstring>
#include <vector>

#include <rocprim/rocprim.hpp>

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

// per window w in [0, W]: first TC sub-block at or after window w
__global__ void k_tc_bounds(const int32_t* __restrict__ t_window_row, int64_t n_tparts,
                            const int32_t* __restrict__ t_row_offset, int64_t W, int32_t* __restrict__ tc_sub_off) {
    const int64_t w = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (w > W) return;
    const int64_t q = lower_bound(t_window_row, n_tparts, (int32_t)(w * 16));
    tc_sub_off[w] = n_tparts ? t_row_offset[q] : 0;
}

// residual entries of row r: its long-part span then its short-part span
struct ResSpans {
    const int32_t *c_row, *c_row_offset, *cs_row, *cs_row_offset;
    int64_t n_cparts, n_csparts;
};

__device__ inline void row_spans(const ResSpans& R, int32_t r, int32_t& l0, int32_t& l1, int32_t& s0, int32_t& s1) {
    const int64_t a = lower_bound(R.c_row, R.n_cparts, r), b = lower_bound(R.c_row, R.n_cparts, r + 1);
    l0 = R.n_cparts ? R.c_row_offset[a] : 0;
    l1 = R.n_cparts ? R.c_row_offset[b] : 0;
    const int64_t c = lower_bound(R.cs_row, R.n_csparts, r), d = lower_bound(R.cs_row, R.n_csparts, r + 1);
    s0 = R.n_csparts ? R.cs_row_offset[c] : 0;
    s1 = R.n_csparts ? R.cs_row_offset[d] : 0;
}

// per row: its long entries [l0, l1) and short entries [s0, s1) in the part
// lists, and their count
__global__ void k_res_row_counts(ResSpans R, int64_t rows, int32_t* __restrict__ cnt, int4* __restrict__ spans) {
    const int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= rows) return;
    int32_t l0, l1, s0, s1;
    row_spans(R, (int32_t)r, l0, l1, s0, s1);
    cnt[r] = (l1 - l0) + (s1 - s0);
    spans[r] = int4{l0, l1, s0, s1};
}

// 32 lanes per part: a part's entries go to its row at the part's place in the
// row's long list (SHORT = 0) or after the long entries (SHORT = 1).  A part
// outside its row's span (part rows not sorted, or out of range) is skipped
// and reported in *bad.
template <int SHORT>
__global__ void k_res_part_fill(const int32_t* __restrict__ prow, const int32_t* __restrict__ poff, int64_t nparts,
                                int64_t rows, const int4* __restrict__ spans, const int32_t* __restrict__ rowptr,
                                const int32_t* __restrict__ column, const void* __restrict__ value,
                                int value_is_f32, int dtype, int64_t ncols, int32_t* __restrict__ cols,
                                uint16_t* __restrict__ vals, const int32_t* __restrict__ src_edge,
                                int32_t* __restrict__ edge, int32_t* __restrict__ bad) {
    const int lane = threadIdx.x & 31;
    const int64_t stride = ((int64_t)gridDim.x * blockDim.x) >> 5;
    for (int64_t pt = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) >> 5; pt < nparts; pt += stride) {
        const int32_t r = prow[pt], b = poff[pt], e = poff[pt + 1];
        if (r < 0 || r >= rows) {
            if (lane == 0) atomicOr(bad, 1);
            continue;
        }
        const int4 sp = spans[r];
        const int32_t lo = SHORT ? sp.z : sp.x, hi = SHORT ? sp.w : sp.y;
        if (b < lo || e > hi || b > e) {
            if (lane == 0) atomicOr(bad, 1);
            continue;
        }
        const int32_t base = rowptr[r] + (SHORT ? (sp.y - sp.x) + (b - sp.z) : (b - sp.x));
        for (int32_t i = lane; i < e - b; i += 32) {
            cols[base + i] = clean_col(column[b + i], ncols);
            vals[base + i] = to_plan_bits(value, b + i, value_is_f32, dtype);
            if (edge) edge[base + i] = src_edge[b + i];
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

// One row's (v0, v1, p0, p1) of a 2:4 group, normalised as block_online.cpp:190-210:
// both hits -> p0 < p1; one hit at q -> zero-valued companion index (1 if q == 0
// else 0); no hit or a missing group -> indices (0, 1) with zero values.
// With s_edge, (e0, e1) are the CSR edges behind (v0, v1) (-1 for a zero filler).
__device__ inline void slot_pair(const float* s_value, const int8_t* s_pos, const int32_t* s_edge, int32_t g,
                                 int32_t g1, int r, int window, float& v0, float& v1, int& p0, int& p1, int& e0,
                                 int& e1) {
    v0 = v1 = 0.f;
    p0 = 0;
    p1 = 1;
    e0 = e1 = -1;
    if (g >= g1 || r >= window) return;
    const int64_t b = ((int64_t)g * window + r) * 2;
    const int q0 = s_pos[b], q1 = s_pos[b + 1];
    const float x0 = s_value[b], x1 = s_value[b + 1];
    const int d0 = s_edge ? s_edge[b] : -1, d1 = s_edge ? s_edge[b + 1] : -1;
    if (q0 >= 0 && q1 >= 0) {
        if (q0 <= q1) { p0 = q0; p1 = q1; v0 = x0; v1 = x1; e0 = d0; e1 = d1; }
        else          { p0 = q1; p1 = q0; v0 = x1; v1 = x0; e0 = d1; e1 = d0; }
    } else if (q0 >= 0 || q1 >= 0) {
        const int q = q0 >= 0 ? q0 : q1;
        const float x = q0 >= 0 ? x0 : x1;
        const int d = q0 >= 0 ? d0 : d1;
        const int dummy = q == 0 ? 1 : 0;
        if (dummy < q) { p0 = dummy; p1 = q; v1 = x; e1 = d; }
        else           { p0 = q; p1 = dummy; v0 = x; e0 = d; }
    }
}

// Thread (tile t, row r), tile t being tile j of its window, fills the row's
// 4 A words, its index word and (row 0) the tile's columns.
__global__ void k_tiles_from_groups(const int32_t* __restrict__ s_offsets, const int32_t* __restrict__ s_column,
                                    const float* __restrict__ s_value, const int8_t* __restrict__ s_pos,
                                    const int32_t* __restrict__ s_edge, const int32_t* __restrict__ tile_off,
                                    int64_t W, int64_t T, int window, int64_t ncols, int dtype,
                                    int32_t* __restrict__ sp_cols, uint32_t* __restrict__ sp_a,
                                    uint16_t* __restrict__ sp_idx, int2* __restrict__ sp_edge) {
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < T * 16; i += (int64_t)gridDim.x * blockDim.x) {
        const int64_t t = i >> 4;
        const int r = (int)(i & 15);
        const int64_t w = lower_bound(tile_off, W + 1, (int32_t)t + 1) - 1;   // last window starting at or before t
        const int32_t g0 = s_offsets[w], g1 = s_offsets[w + 1];
        const int32_t j = (int32_t)(t - tile_off[w]);
        uint32_t idx = 0;
        for (int c = 0; c < 4; ++c) {
            const int32_t g = g0 + 4 * j + c;
            float v0, v1;
            int p0, p1, e0, e1;
            slot_pair(s_value, s_pos, s_edge, g, g1, r, window, v0, v1, p0, p1, e0, e1);
            const uint32_t lo = dtype == TRSPMM_BF16 ? f32_to_bf16_bits(v0) : f32_to_f16_bits(v0);
            const uint32_t hi = dtype == TRSPMM_BF16 ? f32_to_bf16_bits(v1) : f32_to_f16_bits(v1);
            sp_a[(t * 16 + r) * 4 + c] = lo | (hi << 16);
            if (sp_edge) sp_edge[(t * 16 + r) * 4 + c] = make_int2(e0, e1);
            idx |= (uint32_t)((p0 & 3) | ((p1 & 3) << 2)) << (4 * c);
            if (r == 0)
                for (int q = 0; q < 4; ++q)
                    sp_cols[t * 16 + 4 * c + q] = g < g1 ? clean_col(s_column[4 * g + q], ncols) : -1;
        }
        sp_idx[t * 16 + r] = (uint16_t)idx;
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

__host__ __device__ inline int32_t cdiv(int32_t a, int32_t b) { return (a + b - 1) / b; }

// Work items.  A window is split when its tile work exceeds the caps or one of
// its rows holds more residual entries than caps.res_entries; a split window's
// rows are zeroed first and every contribution to them is an atomic add.
//
// Matrix-core residual engine (res_mfma): every window gets matrix-core items
// only (kind 0, or kinds 1/2/3 chunks when split).  Vector-ALU engine: tile
// items for windows with TC/2:4 work, residual items for the rest (kind 0
// store, kind 1 read-modify-write after the tile kernel, kind 2 per-row
// chunks of split windows).
__global__ void k_item_counts(const int32_t* __restrict__ tc_sub_off, const int32_t* __restrict__ tile_off,
                              const int32_t* __restrict__ res_rowptr, int64_t W, SplitCaps caps, int res_mfma,
                              int32_t* __restrict__ n_titems, int32_t* __restrict__ n_ritems,
                              int32_t* __restrict__ is_split) {
    const int64_t w = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (w >= W) return;
    const int32_t S = tc_sub_off[w + 1] - tc_sub_off[w], T = tile_off[w + 1] - tile_off[w];
    int32_t R = 0, rchunks = 0;
    for (int lr = 0; lr < 16; ++lr) {
        const int32_t d = res_rowptr[16 * w + lr + 1] - res_rowptr[16 * w + lr];
        R += d;
        rchunks += cdiv(d, caps.res_entries);
    }
    const bool split = S > caps.tc_sub || T > caps.tiles || R > caps.res_entries;
    const bool tile = S + T > 0;
    if (res_mfma) {
        n_titems[w] = split ? cdiv(S, caps.tc_sub) + cdiv(T, caps.tiles) + cdiv(R, caps.res_entries) : 1;
        n_ritems[w] = 0;
    } else {
        n_titems[w] = !tile ? 0 : (split ? cdiv(S, caps.tc_sub) + cdiv(T, caps.tiles) : 1);
        n_ritems[w] = split ? rchunks : ((R > 0 || !tile) ? 1 : 0);
    }
    is_split[w] = split ? 1 : 0;
}

__global__ void k_fill_items(const int32_t* __restrict__ tc_sub_off, const int32_t* __restrict__ tile_off,
                             const int32_t* __restrict__ res_rowptr, int64_t W, SplitCaps caps, int res_mfma,
                             const int32_t* __restrict__ toff, const int32_t* __restrict__ roff,
                             const int32_t* __restrict__ soff, int4* __restrict__ titems, int4* __restrict__ ritems,
                             int32_t* __restrict__ split_windows) {
    const int64_t w = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (w >= W) return;
    const bool split = soff[w + 1] != soff[w];
    const int32_t s0 = tc_sub_off[w], s1 = tc_sub_off[w + 1], t0 = tile_off[w], t1 = tile_off[w + 1];
    int32_t kt = toff[w], kr = roff[w];
    if (!split) {
        if (toff[w + 1] > kt) titems[kt] = make_int4((int)w, 0, 0, 0);
        // residual item: STORE (no tile work) or read-modify-write (tile kernel wrote the rows)
        if (roff[w + 1] > kr) ritems[kr] = make_int4((int)w, (s1 - s0) + (t1 - t0) > 0 ? 1 : 0, 0, 0);
        return;
    }
    split_windows[soff[w]] = (int32_t)w;
    for (int32_t b = s0; b < s1; b += caps.tc_sub)
        titems[kt++] = make_int4((int)w, 1, b, b + caps.tc_sub < s1 ? b + caps.tc_sub : s1);
    for (int32_t b = t0; b < t1; b += caps.tiles)
        titems[kt++] = make_int4((int)w, 2, b, b + caps.tiles < t1 ? b + caps.tiles : t1);
    if (res_mfma) {
        const int32_t e0 = res_rowptr[16 * w], e1 = res_rowptr[16 * w + 16];
        for (int32_t b = e0; b < e1; b += caps.res_entries)
            titems[kt++] = make_int4((int)w, 3, b, b + caps.res_entries < e1 ? b + caps.res_entries : e1);
        return;
    }
    for (int lr = 0; lr < 16; ++lr) {
        const int32_t e0 = res_rowptr[16 * w + lr], e1 = res_rowptr[16 * w + lr + 1];
        for (int32_t b = e0; b < e1; b += caps.res_entries)
            ritems[kr++] = make_int4((int)w, 2 | (lr << 8), b, b + caps.res_entries < e1 ? b + caps.res_entries : e1);
    }
}

// 8-byte residual records (col | value << 32 | local row << 48), one wave per
// row; an entry whose column is out of range (-1) keeps a zero value, so a
// fragment built from the high words alone contributes nothing for it
__global__ void k_res_records(const int32_t* __restrict__ rowptr, int64_t rows, const int32_t* __restrict__ cols,
                              const uint16_t* __restrict__ vals, uint64_t* __restrict__ rec) {
    const int64_t r = (int64_t)blockIdx.x * (blockDim.x / 64) + threadIdx.x / 64;
    if (r >= rows) return;
    const uint64_t lr = (uint64_t)(r & 15) << 48;
    for (int32_t e = rowptr[r] + (threadIdx.x & 63); e < rowptr[r + 1]; e += 64) {
        const int32_t c = cols[e];
        rec[e] = (uint64_t)(uint32_t)c | ((uint64_t)(c >= 0 ? vals[e] : (uint16_t)0) << 32) | lr;
    }
}

// sort keys of the residual records: (window, column + 1), one wave per row
__global__ void k_res_sort_keys(const int32_t* __restrict__ rowptr, int64_t rows, const int32_t* __restrict__ cols,
                                uint64_t* __restrict__ keys) {
    const int64_t r = (int64_t)blockIdx.x * (blockDim.x / 64) + threadIdx.x / 64;
    if (r >= rows) return;
    const uint64_t w = (uint64_t)(r / 16) << 32;
    for (int32_t e = rowptr[r] + (threadIdx.x & 63); e < rowptr[r + 1]; e += 64)
        keys[e] = w | (uint32_t)(cols[e] + 1);
}

// value of routed-CSR edge e for a refresh: vals[index ? index[e] : e], 0 for
// an empty slot (e < 0) or an index outside [0, n_vals)
struct NewValues {
    const float* vals;
    int64_t n_vals;
    const int32_t* index;
    int dtype;
    __device__ uint16_t bits(int32_t e) const {
        if (e < 0) return 0;
        const int64_t j = index ? (int64_t)index[e] : (int64_t)e;
        const float f = (j >= 0 && j < n_vals) ? vals[j] : 0.f;
        return dtype == TRSPMM_BF16 ? f32_to_bf16_bits(f) : f32_to_f16_bits(f);
    }
};

__global__ void k_refresh_16(const int32_t* __restrict__ edge, int64_t n, NewValues nv, uint16_t* __restrict__ dst) {
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
        dst[i] = nv.bits(edge[i]);
}

__global__ void k_refresh_pairs(const int2* __restrict__ edge, int64_t n, NewValues nv, uint32_t* __restrict__ dst) {
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
        const int2 e = edge[i];
        dst[i] = (uint32_t)nv.bits(e.x) | ((uint32_t)nv.bits(e.y) << 16);
    }
}

// record value bits (32..47) from the edge; an out-of-range column keeps 0
__global__ void k_refresh_records(const int32_t* __restrict__ edge, int64_t n, NewValues nv,
                                  uint64_t* __restrict__ rec) {
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
        const uint64_t r = rec[i];
        const uint64_t v = (int32_t)(uint32_t)r >= 0 ? nv.bits(edge[i]) : 0;
        rec[i] = (r & ~(0xffffull << 32)) | (v << 32);
    }
}

// ---- windows narrower than 16 rows: every stored nonzero becomes a residual
// record.  Key of an entry: row << 32 | its index in the concatenation
// (TC values, 2:4 slots, long residual entries, short residual entries), so
// sorting the keys orders each row's entries in that concatenation order;
// an empty 2:4 slot or a row past the plan gets row = rows (sorted last).
struct NarrowSrc {
    trspmm_meta_view v;
    int64_t rows, ncols, n_t, n_s;
    int dtype;
    // packed 2:4 tiles (trspmm_plan_create_packed), after the residual lists
    const int32_t *p_col, *p_wto;
    const uint32_t *p_a, *p_meta;
    int64_t n_ptiles, n_res;
};

__device__ inline int64_t part_of(const int32_t* off, int64_t parts, int64_t e) {
    return lower_bound(off, parts + 1, (int32_t)(e + 1)) - 1;           // last part starting at or before e
}

__device__ inline void narrow_put(const NarrowSrc& S, int64_t i, int64_t row, int32_t col, uint16_t bits,
                                  uint64_t* keys, int32_t* cols, uint16_t* vals) {
    const bool ok = row >= 0 && row < S.rows;
    keys[i] = ((uint64_t)(ok ? row : S.rows) << 32) | (uint64_t)(uint32_t)i;
    cols[i] = clean_col(col, S.ncols);
    vals[i] = bits;
}

// TC: one thread per sub-block, its bits in the order the values are stored
// (rows t_window_row + half 8 + b / 8, column b % 8)
__global__ void k_narrow_tc(NarrowSrc S, uint64_t* __restrict__ keys, int32_t* __restrict__ cols,
                            uint16_t* __restrict__ vals) {
    const int64_t sb = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (sb >= S.v.n_tsub) return;
    const int64_t p = part_of(S.v.t_row_offset, S.v.n_tparts, sb);
    const int64_t base = S.v.t_window_row[p];
    int64_t vi = S.v.t_block_offset[sb];
    for (int half = 0; half < 2; ++half) {
        const uint64_t bits = (uint64_t)S.v.t_binary[2 * sb + half];
        for (int b = 0; b < 64; ++b) {
            if (!((bits >> b) & 1ull)) continue;
            narrow_put(S, vi, base + half * 8 + b / 8, S.v.t_column[8 * sb + b % 8],
                       to_plan_bits(S.v.t_value, vi, S.v.t_value_is_f32, S.dtype), keys, cols, vals);
            ++vi;
        }
    }
}

// 2:4: one thread per (group, row, slot); the group's window from s_offsets
__global__ void k_narrow_sptc(NarrowSrc S, uint64_t* __restrict__ keys, int32_t* __restrict__ cols,
                              uint16_t* __restrict__ vals) {
    const int64_t b = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (b >= S.n_s) return;
    const int64_t g = b / (2 * S.v.window), lr = (b / 2) % S.v.window;
    const int64_t w = part_of(S.v.s_offsets, S.v.n_windows, g);
    const int q = S.v.s_pos[b];
    const int64_t row = q >= 0 ? w * S.v.window + lr : -1;
    const float x = S.v.s_value[b];
    narrow_put(S, S.n_t + b, row, q >= 0 ? S.v.s_column[4 * g + q] : -1,
               S.dtype == TRSPMM_BF16 ? f32_to_bf16_bits(x) : f32_to_f16_bits(x), keys, cols, vals);
}

// residual lists: one thread per entry (long list, then short list)
__global__ void k_narrow_res(NarrowSrc S, uint64_t* __restrict__ keys, int32_t* __restrict__ cols,
                             uint16_t* __restrict__ vals) {
    const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int64_t nl = S.v.n_cval, n = nl + S.v.n_csval;
    if (e >= n) return;
    const bool lng = e < nl;
    const int64_t k = lng ? e : e - nl;
    const int64_t p = lng ? part_of(S.v.c_row_offset, S.v.n_cparts, k) : part_of(S.v.cs_row_offset, S.v.n_csparts, k);
    const int64_t row = lng ? S.v.c_row[p] : S.v.cs_row[p];
    const int32_t col = lng ? S.v.c_column[k] : S.v.cs_column[k];
    const uint16_t bits = to_plan_bits(lng ? S.v.c_value : S.v.cs_value, k, S.v.c_value_is_f32, S.dtype);
    narrow_put(S, S.n_t + S.n_s + e, row, col, bits, keys, cols, vals);
}

// packed tiles (block_online.cpp:796-841 layout): one thread per (tile, row,
// group), its two slots; tile row r < window is output row w window + r
__global__ void k_narrow_packed(NarrowSrc S, uint64_t* __restrict__ keys, int32_t* __restrict__ cols,
                                uint16_t* __restrict__ vals) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= S.n_ptiles * 64) return;
    const int64_t t = i >> 6;
    const int r = (int)((i >> 2) & 15), c = (int)(i & 3);
    const int64_t w = part_of(S.p_wto, S.v.n_windows, t);
    const uint32_t a = S.p_a[((t * 8 + (r & 7)) * 4 + c) * 2 + (r >> 3)];
    const uint32_t nib = (S.p_meta[t * 8 + (r & 7)] >> (16 * (r >> 3) + 4 * c)) & 0xfu;
    const int64_t row = r < S.v.window ? w * S.v.window + r : -1;
    for (int j = 0; j < 2; ++j) {
        const int q = (int)((nib >> (2 * j)) & 3u);
        const uint16_t h = (uint16_t)(a >> (16 * j));
        const uint16_t bits = S.dtype == TRSPMM_F16 ? h
                                                    : f32_to_bf16_bits((float)__builtin_bit_cast(_Float16, h));
        narrow_put(S, S.n_t + S.n_s + S.n_res + 2 * i + j, row, S.p_col[t * 16 + 4 * c + q], bits, keys, cols, vals);
    }
}

// rowptr[r] = first sorted key of row >= r, r = 0 .. rows
__global__ void k_narrow_rowptr(const uint64_t* __restrict__ keys, int64_t n, int64_t rows,
                                int32_t* __restrict__ rowptr) {
    const int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (r <= rows) rowptr[r] = (int32_t)lower_bound(keys, n, (uint64_t)r << 32);
}

__global__ void k_low_words(const uint64_t* __restrict__ keys, int64_t n, uint32_t* __restrict__ idx) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) idx[i] = (uint32_t)keys[i];
}

__global__ void k_narrow_gather(const uint32_t* __restrict__ idx, int64_t n, const int32_t* __restrict__ cols,
                                const uint16_t* __restrict__ vals, int32_t* __restrict__ out_cols,
                                uint16_t* __restrict__ out_vals) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) {
        out_cols[i] = cols[idx[i]];
        out_vals[i] = vals[idx[i]];
    }
}

inline unsigned blocks_for(int64_t n, int threads = 256) {
    int64_t b = (n + threads - 1) / threads;
    if (b < 1) b = 1;
    if (b > 65535 * 8) b = 65535 * 8;
    return (unsigned)b;
}

}  // namespace

// dense TC rows: (sub-block s, row r) -> its 8 slots as 16-bit values (zeros
// where the occupancy bit is clear), the MFMA A-fragment row of that lane
__global__ void k_tc_dense(const uint64_t* __restrict__ bits, const int32_t* __restrict__ valoff,
                           const uint16_t* __restrict__ vals, int64_t S, uint16_t* __restrict__ dense) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= 16 * S) return;
    const int64_t sb = i / 16;
    const int r = (int)(i % 16);
    const uint64_t lo = bits[2 * sb], hi = bits[2 * sb + 1];
    const uint64_t word = r < 8 ? lo : hi;
    const int sh = (r & 7) * 8;
    const uint32_t b = (uint32_t)(word >> sh) & 0xffu;
    int vi = valoff[sb] + (r < 8 ? 0 : __popcll(lo)) + __popcll(word & ((1ull << sh) - 1ull));
    for (int q = 0; q < 8; ++q) dense[i * 8 + q] = ((b >> q) & 1u) ? vals[vi++] : (uint16_t)0;
}

int Plan::finalize() {
    dev.tc_sub_off = tc_sub_off.as<int32_t>();
    dev.tc_cols = tc_cols.as<int32_t>();
    dev.tc_bits = tc_bits.as<uint64_t>();
    dev.tc_valoff = tc_valoff.as<int32_t>();
    dev.tc_vals = tc_vals.as<uint16_t>();
    dev.tc_dense = tc_dense.as<uint32_t>();
    dev.sp_tile_off = sp_tile_off.as<int32_t>();
    dev.sp_cols = sp_cols.as<int32_t>();
    dev.sp_a = sp_a.as<uint32_t>();
    dev.sp_idx = sp_idx.as<uint16_t>();
    dev.res_rowptr = res_rowptr.as<int32_t>();
    dev.res_cols = res_cols.as<int32_t>();
    dev.res_vals = res_vals.as<uint16_t>();
    dev.res_rec = res_rec.as<uint64_t>();
    dev.titems = titems.as<int4>();
    dev.ritems = ritems.as<int4>();
    dev.split_windows = split_windows.as<int32_t>();
    dev.queue = queue.as<int32_t>();
    return TRSPMM_OK;
}

// Tile work items heaviest first.  An item's cost is its gathered B rows
// (32 per step of 4 TC sub-blocks, 64 per step of 4 2:4 tiles, 32 per step of
// 32 residual entries) plus its store; the items are sorted by the class
// floor(log2(cost)), stably.  The dispatcher deals consecutive workgroups to
// the 8 XCDs round robin, so the heavy items start first and spread over all
// XCDs; in window order, R-MAT's hub windows (mostly multiples of 8) shared
// one XCD.  Within a class the items (in window order) are cut into 8
// contiguous chunks that are interleaved, so each XCD walks a contiguous run
// of windows, whose columns overlap, through its own L2.
static int sort_items_heaviest_first(Plan& P, int64_t n, hipStream_t s) {
    const int64_t W = P.dev.windows;
    std::vector<int4> it((size_t)n), out((size_t)n);     // `it` in window order
    std::vector<int32_t> so((size_t)W + 1), to((size_t)W + 1), ro((size_t)(16 * W + 1));
    TRSPMM_HIP(hipMemcpyAsync(it.data(), P.titems.as<int4>(), 16 * n, hipMemcpyDeviceToHost, s));
    TRSPMM_HIP(hipMemcpyAsync(so.data(), P.tc_sub_off.as<int32_t>(), 4 * (W + 1), hipMemcpyDeviceToHost, s));
    TRSPMM_HIP(hipMemcpyAsync(to.data(), P.sp_tile_off.as<int32_t>(), 4 * (W + 1), hipMemcpyDeviceToHost, s));
    TRSPMM_HIP(hipMemcpyAsync(ro.data(), P.res_rowptr.as<int32_t>(), 4 * (16 * W + 1), hipMemcpyDeviceToHost, s));
    TRSPMM_HIP(hipStreamSynchronize(s));
    std::vector<uint8_t> cls((size_t)n);
    int64_t count[64] = {};
    for (int64_t k = 0; k < n; ++k) {
        const int4 x = it[k];
        int64_t tc = 0, sp = 0, res = 0;
        if (x.y == 0) {
            tc = so[x.x + 1] - so[x.x];
            sp = to[x.x + 1] - to[x.x];
            res = P.dev.res_mfma ? ro[16 * (int64_t)x.x + 16] - ro[16 * (int64_t)x.x] : 0;
        } else if (x.y == 1) {
            tc = x.w - x.z;
        } else if (x.y == 2) {
            sp = x.w - x.z;
        } else {
            res = x.w - x.z;
        }
        const uint64_t cost = 32 * ((tc + 3) / 4) + 64 * ((sp + 3) / 4) + 32 * ((res + 31) / 32) + 16;
        cls[k] = (uint8_t)__builtin_clzll(cost);                 // smaller = heavier
        ++count[cls[k]];
    }
    int64_t start[64];
    for (int64_t c = 0, run = 0; c < 64; ++c) {
        start[c] = run;
        run += count[c];
    }
    std::vector<int4> sorted((size_t)n);
    for (int64_t k = 0; k < n; ++k) sorted[start[cls[k]]++] = it[k];     // start[c] ends at class c's end
    for (int64_t c = 0, a = 0; c < 64; ++c) {
        const int64_t m = count[c], ch = (m + 7) / 8;
        int64_t o = a;
        for (int64_t t = 0; t < ch; ++t)
            for (int64_t x = 0; x < 8; ++x)
                if (x * ch + t < m) out[o++] = sorted[a + x * ch + t];
        a += m;
    }
    TRSPMM_HIP(hipMemcpyAsync(P.titems.as<int4>(), out.data(), 16 * n, hipMemcpyHostToDevice, s));
    TRSPMM_HIP(hipStreamSynchronize(s));
    return TRSPMM_OK;
}

int Plan::build_items(hipStream_t s) {
    StreamArena A(s);      // temporaries (declared first: back to the cache after the stream is idle)
    const int64_t W = dev.windows;
    if (W == 0) {
        dev.n_titems = dev.n_ritems = dev.n_split = 0;
        return finalize();
    }
    DevBuf nt, nr, sp, toff, roff, soff;
    int rc;
    if ((rc = nt.alloc(A, 4 * (W + 1))) || (rc = nr.alloc(A, 4 * (W + 1))) || (rc = sp.alloc(A, 4 * (W + 1))) ||
        (rc = toff.alloc(A, 4 * (W + 1))) || (rc = roff.alloc(A, 4 * (W + 1))) || (rc = soff.alloc(A, 4 * (W + 1))))
        return rc;
    hipLaunchKernelGGL(k_item_counts, dim3(blocks_for(W)), dim3(256), 0, s, tc_sub_off.as<int32_t>(),
                       sp_tile_off.as<int32_t>(), res_rowptr.as<int32_t>(), W, caps, (int)dev.res_mfma,
                       nt.as<int32_t>(), nr.as<int32_t>(), sp.as<int32_t>());
    TRSPMM_HIP(hipGetLastError());
    if ((rc = exclusive_scan<int32_t>(nt.as<int32_t>(), toff.as<int32_t>(), W, s, &A)) ||
        (rc = exclusive_scan<int32_t>(nr.as<int32_t>(), roff.as<int32_t>(), W, s, &A)) ||
        (rc = exclusive_scan<int32_t>(sp.as<int32_t>(), soff.as<int32_t>(), W, s, &A)))
        return rc;
    int32_t tot[3] = {0, 0, 0};
    TRSPMM_HIP(hipMemcpyAsync(&tot[0], toff.as<int32_t>() + W, 4, hipMemcpyDeviceToHost, s));
    TRSPMM_HIP(hipMemcpyAsync(&tot[1], roff.as<int32_t>() + W, 4, hipMemcpyDeviceToHost, s));
    TRSPMM_HIP(hipMemcpyAsync(&tot[2], soff.as<int32_t>() + W, 4, hipMemcpyDeviceToHost, s));
    TRSPMM_HIP(hipStreamSynchronize(s));
    if ((rc = titems.alloc_cached(16 * (int64_t)(tot[0] ? tot[0] : 1))) || (rc = ritems.alloc_cached(16 * (int64_t)(tot[1] ? tot[1] : 1))) ||
        (rc = split_windows.alloc_cached(4 * (int64_t)(tot[2] ? tot[2] : 1))) || (rc = queue.alloc_cached(4 * kQueueInts)))
        return rc;
    TRSPMM_HIP(hipMemsetAsync(queue.as<int32_t>(), 0, 4 * kQueueInts, s));      // the persistent tile queue starts empty
    hipLaunchKernelGGL(k_fill_items, dim3(blocks_for(W)), dim3(256), 0, s, tc_sub_off.as<int32_t>(),
                       sp_tile_off.as<int32_t>(), res_rowptr.as<int32_t>(), W, caps, (int)dev.res_mfma, toff.as<int32_t>(),
                       roff.as<int32_t>(), soff.as<int32_t>(), titems.as<int4>(), ritems.as<int4>(),
                       split_windows.as<int32_t>());
    TRSPMM_HIP(hipGetLastError());
    TRSPMM_HIP(hipStreamSynchronize(s));
    if ((opts.item_order == 0 || opts.item_order == 2) && tot[0] > 1 && (rc = sort_items_heaviest_first(*this, tot[0], s)))
        return rc;
    dev.n_titems = tot[0];
    dev.n_ritems = tot[1];
    dev.n_split = tot[2];
    return finalize();
}

// The matrix-core residual engine's records of each window sorted by column
// (stable: a column's entries stay in row order).  Each record carries its
// local row, so the order inside a window is free; by column, the waves that
// run at the same time on an XCD sweep the columns of their windows in the
// same direction, so the B rows they share meet in that XCD's L2.  A
// gather-only kernel over R-MAT scale 18's residual entries (heaviest-first
// windows, one wave per window) took 0.139 ms with the entries by column
// against 0.172 ms in row order.
static int sort_records_by_column(Plan& P, int64_t rows, int64_t total, hipStream_t s) {
    StreamArena A(s);      // temporaries (declared first: back to the cache after the stream is idle)
    const int64_t W = P.dev.windows;
    int wbits = 1;
    while (wbits < 31 && ((int64_t)1 << wbits) < W) ++wbits;
    DevBuf keys, keys2, rec2, tmp;
    int rc;
    if ((rc = keys.alloc(A, 8 * total)) || (rc = keys2.alloc(A, 8 * total)) || (rc = rec2.alloc(A, 8 * total))) return rc;
    hipLaunchKernelGGL(k_res_sort_keys, dim3(blocks_for(rows, 4)), dim3(256), 0, s, P.res_rowptr.as<int32_t>(), rows,
                       P.res_cols.as<int32_t>(), keys.as<uint64_t>());
    TRSPMM_HIP(hipGetLastError());
    size_t tmp_bytes = 0;
    TRSPMM_HIP(rocprim::radix_sort_pairs(nullptr, tmp_bytes, keys.as<uint64_t>(), keys2.as<uint64_t>(),
                                         P.res_rec.as<uint64_t>(), rec2.as<uint64_t>(), (size_t)total, 0, 32 + wbits, s));
    if ((rc = tmp.alloc(A, tmp_bytes ? tmp_bytes : 1))) return rc;
    TRSPMM_HIP(rocprim::radix_sort_pairs(tmp.p, tmp_bytes, keys.as<uint64_t>(), keys2.as<uint64_t>(),
                                         P.res_rec.as<uint64_t>(), rec2.as<uint64_t>(), (size_t)total, 0, 32 + wbits, s));
    TRSPMM_HIP(hipMemcpyAsync(P.res_rec.p, rec2.p, 8 * total, hipMemcpyDeviceToDevice, s));
    if (P.has_edges) {
        // the same keys through the same stable sort: the edges land beside their records
        if ((rc = P.rec_edge.alloc_cached(4 * total))) return rc;
        size_t eb = 0;
        TRSPMM_HIP(rocprim::radix_sort_pairs(nullptr, eb, keys.as<uint64_t>(), keys2.as<uint64_t>(),
                                             P.res_edge.as<int32_t>(), P.rec_edge.as<int32_t>(), (size_t)total, 0,
                                             32 + wbits, s));
        if (eb > tmp_bytes) {
            tmp_bytes = eb;
            if ((rc = tmp.alloc(A, tmp_bytes))) return rc;
        }
        TRSPMM_HIP(rocprim::radix_sort_pairs(tmp.p, tmp_bytes, keys.as<uint64_t>(), keys2.as<uint64_t>(),
                                             P.res_edge.as<int32_t>(), P.rec_edge.as<int32_t>(), (size_t)total, 0,
                                             32 + wbits, s));
    }
    TRSPMM_HIP(hipStreamSynchronize(s));     // the temporaries are freed on return
    return TRSPMM_OK;
}

// residual row CSR from the long/short part lists
static int build_residual(const trspmm_meta_view& v, int64_t ncols, int dtype, hipStream_t s, Plan& P,
                          const EdgeMaps* em) {
    StreamArena A(s);      // temporaries (declared first: back to the cache after the stream is idle)
    const int64_t rows = 16 * P.dev.windows;
    int rc;
    DevBuf cnt, spans;
    if ((rc = cnt.alloc(A, 4 * (rows + 1))) || (rc = spans.alloc(A, 16 * (rows + 1))) ||
        (rc = P.res_rowptr.alloc_cached(4 * (rows + 1))))
        return rc;
    ResSpans RS{v.c_row, v.c_row_offset, v.cs_row, v.cs_row_offset, v.n_cparts, v.n_csparts};
    if (rows > 0) {
        hipLaunchKernelGGL(k_res_row_counts, dim3(blocks_for(rows)), dim3(256), 0, s, RS, rows, cnt.as<int32_t>(),
                           spans.as<int4>());
        TRSPMM_HIP(hipGetLastError());
    }
    if ((rc = exclusive_scan<int32_t>(rows ? cnt.as<int32_t>() : nullptr, P.res_rowptr.as<int32_t>(), rows, s, &A)))
        return rc;
    int32_t total = 0;
    TRSPMM_HIP(hipMemcpyAsync(&total, P.res_rowptr.as<int32_t>() + rows, 4, hipMemcpyDeviceToHost, s));
    TRSPMM_HIP(hipStreamSynchronize(s));
    TRSPMM_CHECK(total == v.n_cval + v.n_csval, "residual parts must lie inside the windows' rows");
    P.dev.n_rval = total;
    if (total > 0) {
        if ((rc = P.res_cols.alloc_cached(4 * (int64_t)total)) || (rc = P.res_vals.alloc_cached(2 * (int64_t)total))) return rc;
        if (em && (rc = P.res_edge.alloc_cached(4 * (int64_t)total))) return rc;
        DevBuf bad;
        if ((rc = bad.alloc(A, 4))) return rc;
        TRSPMM_HIP(hipMemsetAsync(bad.p, 0, 4, s));
        if (v.n_cparts > 0)
            hipLaunchKernelGGL(k_res_part_fill<0>, dim3(blocks_for(v.n_cparts, 8)), dim3(256), 0, s, v.c_row,
                               v.c_row_offset, v.n_cparts, rows, spans.as<int4>(), P.res_rowptr.as<int32_t>(),
                               v.c_column, v.c_value, v.c_value_is_f32, dtype, ncols, P.res_cols.as<int32_t>(),
                               P.res_vals.as<uint16_t>(), em ? em->c_edge : nullptr,
                               em ? P.res_edge.as<int32_t>() : nullptr, bad.as<int32_t>());
        if (v.n_csparts > 0)
            hipLaunchKernelGGL(k_res_part_fill<1>, dim3(blocks_for(v.n_csparts, 8)), dim3(256), 0, s, v.cs_row,
                               v.cs_row_offset, v.n_csparts, rows, spans.as<int4>(), P.res_rowptr.as<int32_t>(),
                               v.cs_column, v.cs_value, v.c_value_is_f32, dtype, ncols, P.res_cols.as<int32_t>(),
                               P.res_vals.as<uint16_t>(), em ? em->cs_edge : nullptr,
                               em ? P.res_edge.as<int32_t>() : nullptr, bad.as<int32_t>());
        TRSPMM_HIP(hipGetLastError());
        int32_t hbad = 0;
        TRSPMM_HIP(hipMemcpyAsync(&hbad, bad.p, 4, hipMemcpyDeviceToHost, s));
        TRSPMM_HIP(hipStreamSynchronize(s));
        TRSPMM_CHECK(hbad == 0, "residual parts must list rows in non-decreasing order, inside the windows");
        if ((rc = P.res_rec.alloc_cached(8 * (int64_t)total))) return rc;
        hipLaunchKernelGGL(k_res_records, dim3(blocks_for(rows, 4)), dim3(256), 0, s, P.res_rowptr.as<int32_t>(),
                           rows, P.res_cols.as<int32_t>(), P.res_vals.as<uint16_t>(), P.res_rec.as<uint64_t>());
        TRSPMM_HIP(hipGetLastError());
        if (P.opts.res_order == 0 && (rc = sort_records_by_column(P, rows, total, s))) return rc;
    }
    return TRSPMM_OK;
}

// residual engine of a new plan: one-hot fragments on the matrix cores inside
// the tile kernel (default; measured faster on the R-MAT workloads, it saves
// the second kernel and its read-modify-write of C) or the vector-ALU CSR
// kernel, chosen by the plan options, else by TRSPMM_RESIDUAL=valu|mfma in the
// environment when the plan is built
static int residual_on_matrix_cores(const trspmm_plan_opts& o) {
    if (o.residual_engine == TRSPMM_ENGINE_VALU) return 0;
    if (o.residual_engine == TRSPMM_ENGINE_MFMA) return 1;
    const char* e = std::getenv("TRSPMM_RESIDUAL");
    return (e && std::strcmp(e, "valu") == 0) ? 0 : 1;
}

// feature slab of the tile kernel: the plan options, else TRSPMM_SLAB=64|128, else 64
static int tile_slab(const trspmm_plan_opts& o) {
    if (o.slab == 64 || o.slab == 128) return o.slab;
    const char* e = std::getenv("TRSPMM_SLAB");
    return (e && std::atoi(e) == 128) ? 128 : 64;
}

// Windows narrower than 16 rows (window in [1, 16)): the plan runs the routed
// matrix as the metadata stores it, every nonzero (TC bits, kept 2:4 slots,
// residual lists) a residual record of 16-row windows, each row's entries in
// that order.  The reference's parallel op takes any window in (0, 16]
// (mGCN_online.cpp:885); its 2:4 kernel reads rows s_window * window + r
// (mGCNkernel_online.cu:291-299) and its TC kernel rows t_window_row + r.
static int build_plan_narrow(const trspmm_meta_view& v, int64_t m_out, int64_t ncols, int dtype, hipStream_t s,
                             Plan& P, const NarrowSrc* packed = nullptr) {
    StreamArena A(s);      // temporaries (declared first: back to the cache after the stream is idle)
    TRSPMM_CHECK(v.window >= 1 && v.window < 16, "narrow windows: window must be in [1, 16)");
    TRSPMM_CHECK(v.n_windows >= 0 && v.n_windows * v.window >= m_out, "windows must cover the output rows");
    const int64_t W = (std::max<int64_t>(m_out, v.n_windows * v.window) + 15) / 16;
    const int64_t rows = 16 * W;
    TRSPMM_CHECK(rows < ((int64_t)1 << 31), "too many rows");
    P.dev.windows = W;
    P.dev.m_out = m_out;
    P.dev.ncols = ncols;
    P.dev.window = 16;
    P.dev.dtype = dtype;
    P.dev.n_tsub = 0;
    P.dev.n_tval = 0;
    P.dev.n_tiles = 0;
    const int64_t n_t = v.n_tval, n_s = 2 * v.n_groups * v.window, n_r = v.n_cval + v.n_csval;
    const int64_t n_p = packed ? 128 * packed->n_ptiles : 0;
    const int64_t n = n_t + n_s + n_r + n_p;
    TRSPMM_CHECK(n < ((int64_t)1 << 31), "too many nonzeros");
    NarrowSrc S{v, rows, ncols, n_t, n_s, dtype, nullptr, nullptr, nullptr, nullptr, 0, n_r};
    if (packed) {
        S.p_col = packed->p_col;
        S.p_wto = packed->p_wto;
        S.p_a = packed->p_a;
        S.p_meta = packed->p_meta;
        S.n_ptiles = packed->n_ptiles;
    }
    int rc;
    if ((rc = P.tc_sub_off.alloc_cached(4 * (W + 1))) || (rc = P.sp_tile_off.alloc_cached(4 * (W + 1))) ||
        (rc = P.res_rowptr.alloc_cached(4 * (rows + 1))))
        return rc;
    TRSPMM_HIP(hipMemsetAsync(P.tc_sub_off.p, 0, 4 * (W + 1), s));
    TRSPMM_HIP(hipMemsetAsync(P.sp_tile_off.p, 0, 4 * (W + 1), s));
    DevBuf keys, keys2, idx, idx2, cols, vals, tmp;
    const int64_t n1 = n > 0 ? n : 1;
    if ((rc = keys.alloc(A, 8 * n1)) || (rc = keys2.alloc(A, 8 * n1)) || (rc = idx.alloc(A, 4 * n1)) ||
        (rc = idx2.alloc(A, 4 * n1)) || (rc = cols.alloc(A, 4 * n1)) || (rc = vals.alloc(A, 2 * n1)))
        return rc;
    if (v.n_tsub > 0) {
        hipLaunchKernelGGL(k_narrow_tc, dim3(blocks_for(v.n_tsub)), dim3(256), 0, s, S, keys.as<uint64_t>(),
                           cols.as<int32_t>(), vals.as<uint16_t>());
        TRSPMM_HIP(hipGetLastError());
    }
    if (n_s > 0) {
        hipLaunchKernelGGL(k_narrow_sptc, dim3(blocks_for(n_s)), dim3(256), 0, s, S, keys.as<uint64_t>(),
                           cols.as<int32_t>(), vals.as<uint16_t>());
        TRSPMM_HIP(hipGetLastError());
    }
    if (n_r > 0) {
        hipLaunchKernelGGL(k_narrow_res, dim3(blocks_for(n_r)), dim3(256), 0, s, S, keys.as<uint64_t>(),
                           cols.as<int32_t>(), vals.as<uint16_t>());
        TRSPMM_HIP(hipGetLastError());
    }
    if (n_p > 0) {
        hipLaunchKernelGGL(k_narrow_packed, dim3(blocks_for(n_p / 2)), dim3(256), 0, s, S, keys.as<uint64_t>(),
                           cols.as<int32_t>(), vals.as<uint16_t>());
        TRSPMM_HIP(hipGetLastError());
    }
    int total = 0;
    if (n > 0) {
        int rbits = 1;
        while (rbits < 31 && ((int64_t)1 << rbits) <= rows) ++rbits;
        size_t tmp_bytes = 0;
        TRSPMM_HIP(rocprim::radix_sort_keys(nullptr, tmp_bytes, keys.as<uint64_t>(), keys2.as<uint64_t>(), (size_t)n,
                                            0, 32 + rbits, s));
        if ((rc = tmp.alloc(A, tmp_bytes ? tmp_bytes : 1))) return rc;
        TRSPMM_HIP(rocprim::radix_sort_keys(tmp.p, tmp_bytes, keys.as<uint64_t>(), keys2.as<uint64_t>(), (size_t)n,
                                            0, 32 + rbits, s));
    }
    hipLaunchKernelGGL(k_narrow_rowptr, dim3(blocks_for(rows + 1)), dim3(256), 0, s, keys2.as<uint64_t>(), n, rows,
                       P.res_rowptr.as<int32_t>());
    TRSPMM_HIP(hipGetLastError());
    TRSPMM_HIP(hipMemcpyAsync(&total, P.res_rowptr.as<int32_t>() + rows, 4, hipMemcpyDeviceToHost, s));
    TRSPMM_HIP(hipStreamSynchronize(s));
    P.dev.n_rval = total;
    if (total > 0) {
        // the sorted keys' low words are the entries' indices
        hipLaunchKernelGGL(k_low_words, dim3(blocks_for(total)), dim3(256), 0, s, keys2.as<uint64_t>(), (int64_t)total,
                           idx.as<uint32_t>());
        TRSPMM_HIP(hipGetLastError());
        if ((rc = P.res_cols.alloc_cached(4 * (int64_t)total)) || (rc = P.res_vals.alloc_cached(2 * (int64_t)total)) ||
            (rc = P.res_rec.alloc_cached(8 * (int64_t)total)))
            return rc;
        hipLaunchKernelGGL(k_narrow_gather, dim3(blocks_for(total)), dim3(256), 0, s, idx.as<uint32_t>(),
                           (int64_t)total, cols.as<int32_t>(), vals.as<uint16_t>(), P.res_cols.as<int32_t>(),
                           P.res_vals.as<uint16_t>());
        TRSPMM_HIP(hipGetLastError());
        hipLaunchKernelGGL(k_res_records, dim3(blocks_for(rows, 4)), dim3(256), 0, s, P.res_rowptr.as<int32_t>(),
                           rows, P.res_cols.as<int32_t>(), P.res_vals.as<uint16_t>(), P.res_rec.as<uint64_t>());
        TRSPMM_HIP(hipGetLastError());
        if (P.opts.res_order == 0 && (rc = sort_records_by_column(P, rows, total, s))) return rc;
    }
    TRSPMM_HIP(hipStreamSynchronize(s));   // temporaries are freed on return
    return P.build_items(s);
}

// plan row i -> output row: row_order[i] when it is an output row, else -1
__global__ void k_row_map(const int32_t* __restrict__ order, int64_t n_order, int64_t m_out, int64_t n,
                          int32_t* __restrict__ dst) {
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
        const int32_t r = i < n_order ? order[i] : -1;
        dst[i] = (r >= 0 && (int64_t)r < m_out) ? r : -1;
    }
}

int build_plan(const trspmm_meta_view& v, int64_t m_out, int64_t ncols, int dtype, hipStream_t s, Plan& P,
               const EdgeMaps* em, const int32_t* row_order, int64_t order_rows) {
    StreamArena A(s);      // temporaries (declared first: back to the cache after the stream is idle)
    TRSPMM_CHECK(v.window >= 1 && v.window <= 16, "window must be in [1, 16]");
    TRSPMM_CHECK(P.opts.residual_engine >= 0 && P.opts.residual_engine <= TRSPMM_ENGINE_MFMA,
                 "plan options: residual_engine must be TRSPMM_ENGINE_DEFAULT, _VALU or _MFMA");
    TRSPMM_CHECK(P.opts.slab == 0 || P.opts.slab == 64 || P.opts.slab == 128, "plan options: slab must be 0, 64 or 128");
    P.dev.res_mfma = residual_on_matrix_cores(P.opts);
    P.dev.slab = tile_slab(P.opts);
    TRSPMM_CHECK(P.opts.tile_cap >= 0 && P.opts.tc_cap >= 0 && P.opts.res_cap >= 0,
                 "plan options: caps must be >= 0");
    TRSPMM_CHECK(P.opts.item_order >= 0 && P.opts.item_order <= 3, "plan options: item_order must be 0, 1, 2 or 3");
    TRSPMM_CHECK(P.opts.k_hint >= 0 && P.opts.k_hint <= (1 << 24), "plan options: k_hint must be in [0, 2^24]");
    TRSPMM_CHECK(P.opts.res_order == 0 || P.opts.res_order == 1, "plan options: res_order must be 0 or 1");
    TRSPMM_CHECK(P.opts.tile_kernel == 0 || P.opts.tile_kernel == 1, "plan options: tile_kernel must be 0 or 1");
    P.dev.tile_kernel = P.opts.item_order == 2 ? 1 : P.opts.tile_kernel;
    if (P.opts.k_hint > 0) {
        // Residual work items sized for K: a residual entry's gather grows with K,
        // so the entry cap shrinks as 128 / K (R-MAT scale 18: 4096 at K = 128 was
        // 4 % faster than 2048, and 2048 at K = 256 faster than 1024 or 4096);
        // below K = 128 it grows at most twofold.  The 2:4 cap stays: 32 tiles at
        // K = 256 made the exact tiling 20 % slower.
        const int64_t k = std::max<int64_t>(P.opts.k_hint, 64);
        P.caps.res_entries = (int32_t)std::max<int64_t>(256, 4096 * 128 / k);
    }
    if (P.opts.tile_cap) P.caps.tiles = P.opts.tile_cap;
    if (P.opts.tc_cap) P.caps.tc_sub = P.opts.tc_cap;
    if (P.opts.res_cap) P.caps.res_entries = P.opts.res_cap;
    P.dev.item_order = P.opts.item_order;
    TRSPMM_CHECK(dtype == TRSPMM_F16 || dtype == TRSPMM_BF16, "dtype must be TRSPMM_F16 or TRSPMM_BF16");
    TRSPMM_CHECK(m_out > 0 && ncols > 0, "m_out and ncols must be positive");
    if (v.window != 16) return build_plan_narrow(v, m_out, ncols, dtype, s, P);     // (no edge maps)
    TRSPMM_CHECK(v.n_windows >= 0 && v.n_windows * 16 >= m_out, "windows must cover the output rows");
    const int64_t W = v.n_windows;
    P.dev.windows = W;
    P.dev.m_out = m_out;
    P.dev.ncols = ncols;
    P.dev.window = 16;
    P.dev.dtype = dtype;
    P.dev.n_tsub = v.n_tsub;
    P.dev.n_tval = v.n_tval;
    P.has_edges = em != nullptr;
    P.src_nnz = em ? em->nnz : 0;

    int rc;
    if (row_order) {
        // the row map is applied by the deep tile kernel's stores only
        if (P.dev.tile_kernel != 0 || !P.dev.res_mfma || P.dev.item_order == 2) {
            set_error("a reordered routing runs on the deep tile kernel with the matrix-core residual engine "
                      "(plan options tile_kernel 0, residual engine default or MFMA, item_order 0, 1 or 3)");
            return TRSPMM_EUNSUPPORTED;
        }
        if ((rc = P.row_map.alloc_cached(4 * 16 * (W > 0 ? W : 1)))) return rc;
        hipLaunchKernelGGL(k_row_map, dim3(blocks_for(16 * W)), dim3(256), 0, s, row_order, order_rows, m_out, 16 * W,
                           P.row_map.as<int32_t>());
        TRSPMM_HIP(hipGetLastError());
        P.dev.row_map = P.row_map.as<int32_t>();
    }
    DevBuf tile_cnt;
    if ((rc = P.tc_sub_off.alloc_cached(4 * (W + 1))) || (rc = P.sp_tile_off.alloc_cached(4 * (W + 1))) ||
        (rc = tile_cnt.alloc(A, 4 * (W + 1))))
        return rc;
    hipLaunchKernelGGL(k_tc_bounds, dim3(blocks_for(W + 1)), dim3(256), 0, s, v.t_window_row, v.n_tparts,
                       v.t_row_offset, W, P.tc_sub_off.as<int32_t>());
    TRSPMM_HIP(hipGetLastError());

    // TC sub-blocks
    if (v.n_tsub > 0) {
        if ((rc = P.tc_cols.alloc_cached(4 * 8 * v.n_tsub)) || (rc = P.tc_bits.alloc_cached(8 * 2 * v.n_tsub)) ||
            (rc = P.tc_valoff.alloc_cached(4 * (v.n_tsub + 1))) || (rc = P.tc_vals.alloc_cached(2 * (v.n_tval > 0 ? v.n_tval : 1))))
            return rc;
        hipLaunchKernelGGL(k_copy_cols, dim3(blocks_for(8 * v.n_tsub)), dim3(256), 0, s, v.t_column, 8 * v.n_tsub,
                           ncols, P.tc_cols.as<int32_t>());
        TRSPMM_HIP(hipMemcpyAsync(P.tc_bits.p, v.t_binary, 16 * v.n_tsub, hipMemcpyDeviceToDevice, s));
        TRSPMM_HIP(hipMemcpyAsync(P.tc_valoff.p, v.t_block_offset, 4 * (v.n_tsub + 1), hipMemcpyDeviceToDevice, s));
        if (v.n_tval > 0)
            hipLaunchKernelGGL(k_copy_vals, dim3(blocks_for(v.n_tval)), dim3(256), 0, s, v.t_value, v.n_tval,
                               v.t_value_is_f32, dtype, P.tc_vals.as<uint16_t>());
        TRSPMM_HIP(hipGetLastError());
        if (em && v.n_tval > 0) {
            if ((rc = P.tc_edge.alloc_cached(4 * v.n_tval))) return rc;
            TRSPMM_HIP(hipMemcpyAsync(P.tc_edge.p, em->t_edge, 4 * v.n_tval, hipMemcpyDeviceToDevice, s));
        }
        if ((rc = P.tc_dense.alloc_cached(2 * 16 * 8 * v.n_tsub))) return rc;
        hipLaunchKernelGGL(k_tc_dense, dim3(blocks_for(16 * v.n_tsub)), dim3(256), 0, s, P.tc_bits.as<uint64_t>(),
                           P.tc_valoff.as<int32_t>(), P.tc_vals.as<uint16_t>(), v.n_tsub, P.tc_dense.as<uint16_t>());
        TRSPMM_HIP(hipGetLastError());
    }

    // residual route: one CSR over the padded rows
    if ((rc = build_residual(v, ncols, dtype, s, P, em))) return rc;

    // 2:4 tiles
    if (W > 0) {
        hipLaunchKernelGGL(k_tile_counts, dim3(blocks_for(W)), dim3(256), 0, s, v.s_offsets, W, tile_cnt.as<int32_t>());
        TRSPMM_HIP(hipGetLastError());
    }
    if ((rc = exclusive_scan<int32_t>(W > 0 ? tile_cnt.as<int32_t>() : nullptr, P.sp_tile_off.as<int32_t>(), W, s, &A)))
        return rc;
    int32_t n_tiles = 0;
    TRSPMM_HIP(hipMemcpyAsync(&n_tiles, P.sp_tile_off.as<int32_t>() + W, 4, hipMemcpyDeviceToHost, s));
    TRSPMM_HIP(hipStreamSynchronize(s));
    P.dev.n_tiles = n_tiles;
    if (n_tiles > 0) {
        if ((rc = P.sp_cols.alloc_cached(4 * 16 * (int64_t)n_tiles)) || (rc = P.sp_a.alloc_cached(4 * 64 * (int64_t)n_tiles)) ||
            (rc = P.sp_idx.alloc_cached(2 * 16 * (int64_t)n_tiles)))
            return rc;
        if (em && (rc = P.sp_edge.alloc_cached(8 * 64 * (int64_t)n_tiles))) return rc;
        hipLaunchKernelGGL(k_tiles_from_groups, dim3(blocks_for(16 * (int64_t)n_tiles)), dim3(256), 0, s,
                           v.s_offsets, v.s_column, v.s_value, v.s_pos, em ? em->s_edge : nullptr,
                           P.sp_tile_off.as<int32_t>(), W, (int64_t)n_tiles, v.window, ncols, dtype,
                           P.sp_cols.as<int32_t>(), P.sp_a.as<uint32_t>(), P.sp_idx.as<uint16_t>(),
                           em ? P.sp_edge.as<int2>() : nullptr);
        TRSPMM_HIP(hipGetLastError());
    }
    TRSPMM_HIP(hipStreamSynchronize(s));   // temporaries are freed on return
    return P.build_items(s);
}

int build_plan_packed(const int32_t* s_tile_column, const int32_t* s_packed_a, const int32_t* s_packed_meta,
                      const int32_t* s_window_tile_offset, int64_t T, int64_t W, int window, int64_t m_out,
                      int64_t ncols, hipStream_t s, Plan& P) {
    TRSPMM_CHECK(window >= 1 && window <= 16, "window must be in [1, 16]");
    if (window != 16) {
        // narrower windows: the tiles' rows w window + r (r < window) as residual records
        TRSPMM_CHECK(m_out > 0 && ncols > 0, "m_out and ncols must be positive");
        trspmm_meta_view v{};
        v.window = window;
        v.n_windows = W;
        NarrowSrc pk{};
        pk.p_col = s_tile_column;
        pk.p_wto = s_window_tile_offset;
        pk.p_a = (const uint32_t*)s_packed_a;
        pk.p_meta = (const uint32_t*)s_packed_meta;
        pk.n_ptiles = T;
        P.dev.res_mfma = residual_on_matrix_cores(P.opts);
        P.dev.slab = tile_slab(P.opts);
        return build_plan_narrow(v, m_out, ncols, TRSPMM_F16, s, P, &pk);
    }
    TRSPMM_CHECK(m_out > 0 && ncols > 0, "m_out and ncols must be positive");
    TRSPMM_CHECK(W * 16 >= m_out, "windows must cover the output rows");
    P.dev.windows = W;
    P.dev.m_out = m_out;
    P.dev.ncols = ncols;
    P.dev.dtype = TRSPMM_F16;
    P.dev.n_tiles = T;
    int rc;
    if ((rc = P.tc_sub_off.alloc_cached(4 * (W + 1))) || (rc = P.res_rowptr.alloc_cached(4 * (16 * W + 1))) ||
        (rc = P.sp_tile_off.alloc_cached(4 * (W + 1))))
        return rc;
    TRSPMM_HIP(hipMemsetAsync(P.tc_sub_off.p, 0, 4 * (W + 1), s));
    TRSPMM_HIP(hipMemsetAsync(P.res_rowptr.p, 0, 4 * (16 * W + 1), s));
    TRSPMM_HIP(hipMemcpyAsync(P.sp_tile_off.p, s_window_tile_offset, 4 * (W + 1), hipMemcpyDeviceToDevice, s));
    if (T > 0) {
        if ((rc = P.sp_cols.alloc_cached(4 * 16 * T)) || (rc = P.sp_a.alloc_cached(4 * 64 * T)) || (rc = P.sp_idx.alloc_cached(2 * 16 * T)))
            return rc;
        hipLaunchKernelGGL(k_tiles_from_packed, dim3(blocks_for(T * 64)), dim3(256), 0, s, s_tile_column,
                           (const uint32_t*)s_packed_a, (const uint32_t*)s_packed_meta, T, ncols,
                           P.sp_cols.as<int32_t>(), P.sp_a.as<uint32_t>(), P.sp_idx.as<uint16_t>());
        TRSPMM_HIP(hipGetLastError());
    }
    TRSPMM_HIP(hipStreamSynchronize(s));
    return P.build_items(s);
}

// Value refresh: the routes, tiles, records and work items stay; only the
// 16-bit values are rewritten, each slot from the CSR edge it was built from,
// in the plan's own layouts (the dense TC rows are rebuilt from the new TC
// values).  Stream-ordered and allocation-free.
int plan_set_values(Plan& P, const float* vals, int64_t n_vals, const int32_t* index, int64_t nnz, hipStream_t s) {
    if (!P.has_edges) {
        set_error("plan_set_values: the plan carries no edge maps (route with TRSPMM_ROUTE_EDGES and build it "
                  "with trspmm_plan_from_meta)");
        return TRSPMM_EUNSUPPORTED;
    }
    TRSPMM_CHECK(nnz == P.src_nnz, "plan_set_values: nnz must be the routed CSR's nonzero count");
    TRSPMM_CHECK(n_vals >= 0, "plan_set_values: n_vals must be >= 0");
    TRSPMM_CHECK(n_vals == 0 || vals, "plan_set_values: vals must not be null");
    const NewValues nv{vals, n_vals, index, P.dev.dtype};
    const int64_t S = P.dev.n_tsub, TV = P.dev.n_tval, T = P.dev.n_tiles, R = P.dev.n_rval;
    if (TV > 0) {
        hipLaunchKernelGGL(k_refresh_16, dim3(blocks_for(TV)), dim3(256), 0, s, P.tc_edge.as<int32_t>(), TV, nv,
                           P.tc_vals.as<uint16_t>());
        TRSPMM_HIP(hipGetLastError());
    }
    if (S > 0) {
        hipLaunchKernelGGL(k_tc_dense, dim3(blocks_for(16 * S)), dim3(256), 0, s, P.tc_bits.as<uint64_t>(),
                           P.tc_valoff.as<int32_t>(), P.tc_vals.as<uint16_t>(), S, P.tc_dense.as<uint16_t>());
        TRSPMM_HIP(hipGetLastError());
    }
    if (T > 0) {
        hipLaunchKernelGGL(k_refresh_pairs, dim3(blocks_for(64 * T)), dim3(256), 0, s, P.sp_edge.as<int2>(), 64 * T,
                           nv, P.sp_a.as<uint32_t>());
        TRSPMM_HIP(hipGetLastError());
    }
    if (R > 0) {
        hipLaunchKernelGGL(k_refresh_16, dim3(blocks_for(R)), dim3(256), 0, s, P.res_edge.as<int32_t>(), R, nv,
                           P.res_vals.as<uint16_t>());
        TRSPMM_HIP(hipGetLastError());
        const int32_t* re = P.rec_edge.p ? P.rec_edge.as<int32_t>() : P.res_edge.as<int32_t>();
        hipLaunchKernelGGL(k_refresh_records, dim3(blocks_for(R)), dim3(256), 0, s, re, R, nv,
                           P.res_rec.as<uint64_t>());
        TRSPMM_HIP(hipGetLastError());
    }
    return TRSPMM_OK;
}

}  // namespace trspmm
