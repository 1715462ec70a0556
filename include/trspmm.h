/*
 * trspmm — three-way routed SpMM for AMD Instinct MI355X (gfx950).
 *
 * C ABI of libtrspmm.so.  Plain pointers and sizes only; device pointers are
 * HIP device memory, `stream` is a hipStream_t passed as void* (NULL = the
 * null stream).  Every function returns 0 on success or a TRSPMM_E* code and
 * leaves a message for trspmm_last_error() (per calling thread).  Calls are
 * stream-ordered; only functions documented as synchronous block the host.
 *
 * Reference interfaces replaced (paths relative to the TR-SPMM sources):
 *   trspmm_route          <- Libra5BlockOnline.block_sptc_2to4_online
 *                            TR-source/Block/block_online.cpp:545-549, :873-876
 *   trspmm_plan_create    <- the device upload + argument contract of
 *                            spmm_forward_fp16_tcu_cuda_sptc_mma_parallel_online
 *                            TR-source/SpMM/mGCN_online.cpp:760-1042
 *   trspmm_plan_create_packed
 *                         <- spmm_forward_fp16_sptc_mma_packed_online inputs
 *                            TR-source/SpMM/mGCN_online.cpp:654-706
 *   trspmm_spmm           <- spmm_forward_fp16_tcu_cuda_sptc_mma_parallel_online_kernel
 *                            TR-source/SpMM/mGCNkernel_online.cu:1529-1727 (one launch,
 *                            single-owner output rows instead of 4 streams + add)
 *   trspmm_spmm_timed     <- the warm-up / timed-epoch loop of the same launcher
 *                            (mGCNkernel_online.cu:1604-1704) and its ms outputs
 *   trspmm_tca_order, trspmm_minhash
 *                         <- apply_tca_reorder, spmm/tr_csr_fp16/tr_tca_reorder.py:44-357
 *   trspmm_match_2to4, _hash, _bucket, _online
 *                         <- the setup_24matching statistics ops
 *                            matching_utils{,_hash,_bucket,_online}.match_2to4*
 *                            (TR-source/Block/block_matching_*.cpp / .cu)
 * Entry points with no reference counterpart (the reference has no backward
 * and no device router):
 *   trspmm_csr_transpose, trspmm_sddmm, trspmm_bwd_plan_*, trspmm_spmm_t_sddmm
 *   (backward pass), trspmm_plan_*_ex (plan options), trspmm_plan_set_values
 *   (value refresh), trspmm_memcpy,
 *   trspmm_host_sort_order (diagnostics).
 */
#ifndef TRSPMM_H
#define TRSPMM_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum {
    TRSPMM_OK = 0,
    TRSPMM_EINVAL = 1,       /* bad argument (shape, range, dtype, ordering) */
    TRSPMM_EHIP = 2,         /* a HIP runtime call failed */
    TRSPMM_ENOMEM = 3,       /* device allocation failed */
    TRSPMM_EUNSUPPORTED = 4  /* valid request this build does not implement */
};

enum { TRSPMM_F16 = 0, TRSPMM_BF16 = 1 };

/* route masks for trspmm_spmm: which routes of the plan contribute */
enum { TRSPMM_ROUTE_TC = 1, TRSPMM_ROUTE_SPTC = 2, TRSPMM_ROUTE_RESIDUAL = 4, TRSPMM_ROUTE_ALL = 7 };

/* router output modes */
enum {
    TRSPMM_MODE_COMPAT = 0,  /* bit-exact with the reference (3rd/4th hits of a 2:4 row dropped) */
    TRSPMM_MODE_EXACT = 1,   /* t_* / s_* bit-exact; dropped hits appended to the residual route */
    TRSPMM_MODE_TUNED = 2    /* wave64 thresholds: low-occupancy 2:4 groups go to the residual route */
};

const char* trspmm_last_error(void);
int trspmm_version(void);
int trspmm_device_arch(char* buf, int len);   /* gcnArchName of the current device */
/* No counterpart in the reference.  Router results and router scratch take
 * their device blocks from a per-device cache that keeps released blocks for
 * the next call (hipMalloc / hipFree cost ~0.3 ms each and hipFree waits for
 * the device); at most TRSPMM_CACHE_MB (environment, default 4096) idle MiB per
 * device are kept.  This frees every idle block and stores the bytes freed in
 * *freed (may be NULL). */
int trspmm_cache_trim(int64_t* freed);

/* ------------------------------------------------------------------ router */

typedef struct trspmm_route_cfg {
    int32_t part_t;         /* partSize_t: TC sub-blocks per part (32) */
    int32_t part_c;         /* partSize_c: residual nonzeros per long part (32) */
    int32_t short_c;        /* shortSize: residual parts of <= short_c go to the short list (3) */
    int32_t density;        /* TC column threshold within a window (8) */
    int32_t window;         /* rows per window, 1..16 (16) */
    int32_t wide;           /* range-checked only, 1..32, as the reference (16) */
    float tc_min_util;      /* TC chunk utilisation floor (0.0) */
    int32_t skip_tc;        /* LIBRA_SKIP_TC */
    int32_t unit_min_nnz;   /* LIBRA_SPTC_UNIT_MIN_NNZ (2) */
    int32_t mode;           /* TRSPMM_MODE_* */
    int32_t sptc_min_nnz;   /* TUNED: a 2:4 group with fewer nonzeros goes to the residual route
                               (0 keeps every group, 33 none) */
    int32_t flags;          /* TRSPMM_ROUTE_* bits (0 = none) */
} trspmm_route_cfg;

/* trspmm_route_cfg.flags: also record the CSR edge behind every value slot
 * (no reference counterpart).  Plans built from such a router result carry
 * the maps and accept new values with trspmm_plan_set_values; the 29 arrays
 * are unchanged. */
enum { TRSPMM_ROUTE_EDGES = 1, TRSPMM_ROUTE_REORDER = 2 };
/* TRSPMM_ROUTE_REORDER (EXACT or TUNED mode, window 16; no reference
 * counterpart, in the spirit of the reference caller's row reordering,
 * spmm/tr_csr_fp16/tr_tca_reorder.py): the router also routes the rows in
 * order of (first column, nonzeros descending, row index), empty rows last,
 * and keeps that routing when its routes gather at least a tenth fewer B
 * rows per multiply (TC sub-block columns + 2:4 tile columns + residual
 * entries).  The arrays then describe the reordered rows,
 * trspmm_meta_row_order gives the order, edge maps still name the input's
 * edges, and plans built from the result write every row to its input
 * index. */

typedef struct trspmm_meta trspmm_meta;

/* Routes a canonical CSR (each row's columns strictly ascending) that lives
 * in device memory.  `rows` is the (window-padded) row count, `ncols` the
 * column count (rows of B).  Produces the reference's 29 metadata arrays on
 * the device (index order of block_online.cpp:844-868).  Synchronous: the
 * array sizes are known on the host when it returns. */
int trspmm_route(const int32_t* d_rowptr, const int32_t* d_col, const float* d_val, int64_t rows,
                 int64_t nnz, int64_t ncols, const trspmm_route_cfg* cfg, void* stream,
                 trspmm_meta** out);
int64_t trspmm_meta_count(const trspmm_meta* meta, int index);  /* elements of array `index` */
void* trspmm_meta_data(const trspmm_meta* meta, int index);       /* device pointer */
double trspmm_meta_seconds(const trspmm_meta* meta);             /* router wall time */
/* Device pointer to the routed rows' input indices (routed row i is input row
 * order[i], `rows` entries) when the result kept a reordered routing
 * (TRSPMM_ROUTE_REORDER); NULL when the rows kept their order.  No reference
 * counterpart. */
const int32_t* trspmm_meta_row_order(const trspmm_meta* meta);
void trspmm_meta_free(trspmm_meta* meta);

/* ------------------------------------------------------------------ plans */

/* Device views of the reference-format metadata (the op's inputs,
 * mGCN_online.cpp:760-800).  Value arrays may be fp32 (router output) or
 * fp16 (what the reference's caller passes, mdataset2_online.py:151-153);
 * `*_value_is_f32` says which. */
typedef struct trspmm_meta_view {
    const int32_t* t_row_offset;   int64_t n_tparts;      /* [n_tparts + 1] */
    const int32_t* t_block_offset;                         /* [n_tsub + 1]   */
    const int32_t* t_column;       int64_t n_tsub;         /* [8 n_tsub]     */
    const void* t_value;           int64_t n_tval;
    const int32_t* t_window_row;                           /* [n_tparts]     */
    const int64_t* t_binary;                               /* [2 n_tsub]     */
    const int32_t* c_row_offset;   int64_t n_cparts;
    const int32_t* c_row;
    const int32_t* c_column;       int64_t n_cval;
    const void* c_value;
    const int32_t* cs_row_offset;  int64_t n_csparts;
    const int32_t* cs_row;
    const int32_t* cs_column;      int64_t n_csval;
    const void* cs_value;
    const int32_t* s_column;       int64_t n_groups;       /* [n_groups, 4]  */
    const float* s_value;                                  /* [n_groups, window, 2] */
    const int8_t* s_pos;                                   /* [n_groups, window, 2] */
    const int32_t* s_offsets;      int64_t n_windows;      /* [n_windows + 1] */
    int32_t t_value_is_f32, c_value_is_f32;
    int32_t window;
    int32_t reserved;
} trspmm_meta_view;

typedef struct trspmm_plan trspmm_plan;

/* Plan options (no reference counterpart; the reference's kernels are fixed).
 * Zero-initialised = defaults (environment TRSPMM_RESIDUAL / TRSPMM_SLAB, else
 * the matrix-core residual engine and 64-feature slabs; caps 64 / 64 / 2048;
 * heaviest-first work items dealt round robin). */
enum { TRSPMM_ENGINE_DEFAULT = 0, TRSPMM_ENGINE_VALU = 1, TRSPMM_ENGINE_MFMA = 2 };
typedef struct trspmm_plan_opts {
    int32_t residual_engine;  /* TRSPMM_ENGINE_* */
    int32_t slab;             /* tile-kernel features per pass: 0 (default), 64 or 128 */
    int32_t tile_cap;         /* a window with more 2:4 tiles is split into chunks of this many (0 = 64) */
    int32_t tc_cap;           /* the same for TC sub-blocks (0 = 64) */
    int32_t res_cap;          /* the same for residual entries (0 = 2048, or sized by k_hint) */
    int32_t item_order;       /* 0: one workgroup per work item, the items sorted heaviest first
                                    (by gathered rows, in classes of powers of two) and dealt
                                    round robin to the XCDs (the dispatch order);
                                 1: one workgroup per work item in window order, each XCD takes
                                    a contiguous range;
                                 2: persistent tile kernel, a chip-sized grid whose workgroups
                                    take the heaviest-first items from the plan's queue counters
                                    (launches of one such plan must not overlap on concurrent
                                    streams);
                                 3: as 0, with the items in window order */
    int32_t k_hint;           /* feature width K the plan will mostly run at (0 = not known).
                                 When > 0 and res_cap is 0, the residual cap is sized for it,
                                 so that a residual work item holds about the same work at
                                 any K: 4096 * 128 / K' entries (at least 256),
                                 K' = max(K, 64).  Any K still runs on the plan. */
    int32_t res_order;        /* order of the matrix-core residual engine's entries inside a
                                 window: 0 = by column (the default; concurrent windows then
                                 sweep their columns in the same direction and share B rows
                                 in L2), 1 = row order (the residual CSR's) */
    int32_t tile_kernel;      /* 0 = the deep-pipelined tile kernel (the default: full 128-feature
                                 B rows per step, 16 KB of gathers in flight per wave);
                                 1 = the earlier kernel (64-feature slabs, one sub-step ahead),
                                 kept for comparison.  Persistent plans (item_order 2) run 1. */
    int32_t reserved;
} trspmm_plan_opts;

/* Builds the device plan (AMD lane layouts) from reference-format metadata.
 * m_out: output rows (mOri); ncols: rows of B (kOri); dtype: TRSPMM_F16/BF16.
 * view->window in [1, 16]: at 16 the routes keep their tiles; a narrower
 * window's nonzeros run as residual records of 16-row windows (2:4 rows
 * window_index * window + r, TC rows t_window_row + r, as the reference's
 * kernels read them: mGCNkernel_online.cu:291-299). */
int trspmm_plan_create(const trspmm_meta_view* view, int64_t m_out, int64_t ncols, int dtype,
                       void* stream, trspmm_plan** out);
/* Same, from the router's own output. */
int trspmm_plan_from_meta(const trspmm_meta* meta, int64_t m_out, int64_t ncols, int dtype,
                          void* stream, trspmm_plan** out);
/* The same two, with plan options (opts may be null = defaults). */
int trspmm_plan_create_ex(const trspmm_meta_view* view, int64_t m_out, int64_t ncols, int dtype,
                          const trspmm_plan_opts* opts, void* stream, trspmm_plan** out);
int trspmm_plan_from_meta_ex(const trspmm_meta* meta, int64_t m_out, int64_t ncols, int dtype,
                             const trspmm_plan_opts* opts, void* stream, trspmm_plan** out);
/* 2:4 route only, from the mma.sp packed tiles (fp16 values); tile row r of
 * window w is output row w * window + r (r < window), window in [1, 16]. */
int trspmm_plan_create_packed(const int32_t* s_tile_column, const int32_t* s_packed_a,
                              const int32_t* s_packed_meta, const int32_t* s_window_tile_offset,
                              int64_t n_tiles, int64_t n_windows, int32_t window, int64_t m_out,
                              int64_t ncols, void* stream, trspmm_plan** out);
void trspmm_plan_free(trspmm_plan* plan);
/* New values for a plan built from a router result with TRSPMM_ROUTE_EDGES
 * (no reference counterpart; the reference's op takes its values on every
 * call, mGCN_online.cpp:760-800, which this makes possible without routing
 * again).  The value of the routed CSR's edge e becomes
 * vals[index ? index[e] : e] (fp32, device memory), converted to the plan's
 * operand type exactly as at plan creation; index (device, nnz entries, may
 * be null) serves plans of A^T fed with A's values (index = the perm of
 * trspmm_csr_transpose).  nnz = the routed CSR's nonzeros; n_vals = entries of
 * vals (reads outside [0, n_vals) give 0).  Stream-ordered, no allocation;
 * must not overlap a multiply on the same plan.  TRSPMM_EUNSUPPORTED for a
 * plan without edge maps. */
int trspmm_plan_set_values(trspmm_plan* plan, const float* vals, int64_t n_vals, const int32_t* index, int64_t nnz,
                           void* stream);
/* counts: [windows, tc sub-blocks, tc nnz, 2:4 tiles, residual work items, residual nnz] */
int trspmm_plan_stats(const trspmm_plan* plan, int64_t* counts6);

/* ------------------------------------------------------------------ multiply */

/* C[m_out, K] (ldc) = A_routed · B[ncols, K] (ldb), FP32 accumulate, fp32
 * output, every output row written once.  B dtype = plan dtype.  K % 8 == 0,
 * ldb % 8 == 0, ldc % 4 == 0, B and C 16-byte aligned. */
int trspmm_spmm(const trspmm_plan* plan, const void* B, int64_t ldb, int64_t K, float* C,
                int64_t ldc, uint32_t route_mask, void* stream);

/* The reference's timing loop: `warmup` untimed launches, then `iters` timed
 * ones with hipEvents; *ms_avg = mean ms per launch.  Synchronous. */
int trspmm_spmm_timed(const trspmm_plan* plan, const void* B, int64_t ldb, int64_t K, float* C,
                      int64_t ldc, uint32_t route_mask, int warmup, int iters, void* stream,
                      float* ms_avg);

/* Copies `bytes` between any host/device addresses, ordered on `stream`;
 * returns after the copy completed (used to hand router output to callers). */
int trspmm_memcpy(void* dst, const void* src, int64_t bytes, void* stream);

/* ------------------------------------------------------------------ backward
 * No counterpart in the reference (its SpMM has no backward). */

/* A^T of a canonical CSR A (rows x ncols), on the device: t_rowptr[ncols + 1],
 * t_col[nnz] (row indices of A, ascending within each column), t_val[nnz]
 * (may be null) and perm[nnz] (A's edge index of each A^T entry), all
 * caller-allocated device memory.  Synchronous (returns after the stream
 * finished).  Rejects column indices outside [0, ncols). */
int trspmm_csr_transpose(const int32_t* d_rowptr, const int32_t* d_col, const float* d_val, int64_t rows,
                         int64_t nnz, int64_t ncols, int32_t* t_rowptr, int32_t* t_col, float* t_val,
                         int32_t* perm, void* stream);

/* dvals[e] = sum_f X[row_e, f] * Y[col_e, f] for every edge e of the CSR
 * (rowptr[rows + 1], col[nnz]; device memory; Y has ncols rows, and an edge
 * whose column lies outside [0, ncols) gets 0 without Y being read): the gradient of
 * sum(dY * (A·B)) with respect to A's values, X = dY [rows, K] (ldx),
 * Y = B [ncols, K] (ldy), both `dtype` (TRSPMM_F16 / TRSPMM_BF16), fp32
 * accumulation, fp32 dvals in A's edge order.  K, ldx, ldy multiples of 8;
 * X and Y 16-byte aligned.  Stream-ordered. */
int trspmm_sddmm(const int32_t* d_rowptr, int64_t rows, const int32_t* d_col, int64_t nnz, int64_t ncols,
                 const void* X, int64_t ldx, const void* Y, int64_t ldy, int64_t K, int dtype, float* dvals,
                 void* stream);

/* Fused backward over the rows of A^T: one pass computes both
 *   dB[N, K]  = A^T · dY          (fp32, ldd: every row written)
 *   dvals[e]  = <dY[row_e], B[col_e]>   (fp32, A's edge order)
 * gathering each dY row once for both: A^T is routed into TC sub-blocks
 * (every column two rows of a window share) and residual records, its rows
 * reordered when that gathers fewer dY rows.  The plan is built once from A^T as
 * trspmm_csr_transpose returns it (t_rowptr[rows + 1], t_col, t_val, perm;
 * device memory; rows = ncols of A, ncols = rows of A) and holds its own
 * copies, the values in `dtype`, and a scratch array of nnz floats: calls on
 * one plan must not overlap.  K multiple of 8, at most 512; dY [ncols, K]
 * (lddy) and B [rows, K] (ldb) in `dtype`, 16-byte aligned. */
typedef struct trspmm_bwd_plan trspmm_bwd_plan;
/* Rejects (TRSPMM_EINVAL) column indices outside [0, ncols) and a perm that
 * is not a permutation of [0, nnz). */
int trspmm_bwd_plan_create(const int32_t* t_rowptr, const int32_t* t_col, const float* t_val,
                           const int32_t* perm, int64_t rows, int64_t nnz, int64_t ncols, int dtype,
                           void* stream, trspmm_bwd_plan** out);
void trspmm_bwd_plan_free(trspmm_bwd_plan* plan);
/* New values of A (vals [nnz], fp32, A's edge order, device memory) for a
 * fused-backward plan; stream-ordered, no allocation, must not overlap a call
 * on the same plan.  No reference counterpart. */
int trspmm_bwd_plan_set_values(trspmm_bwd_plan* plan, const float* vals, int64_t nnz, void* stream);
/* What the fused backward runs on (no reference counterpart): counts5 =
 * [1 if the routed fused kernel runs (else the lane-group kernels), A^T's TC
 * sub-blocks, TC nonzeros, residual records, 1 if A^T's rows were reordered]. */
int trspmm_bwd_plan_stats(const trspmm_bwd_plan* plan, int64_t* counts5);
int trspmm_spmm_t_sddmm(const trspmm_bwd_plan* plan, const void* dY, int64_t lddy, const void* B,
                        int64_t ldb, int64_t K, float* dB, int64_t lddb, float* dvals, void* stream);

/* ------------------------------------------------------------------ matching statistics
 * Replaces matching_utils.match_2to4 (setup_24matching op,
 * TR-source/Block/block_matching_window_para_dr.cpp:346-419): statistics of
 * the sliding-window 1:2 / 2:4 matching heuristic with dense-pool routing
 * over the 16-bit row masks of each window.  Host memory, host threads
 * (threads <= 0: up to 16).  values are not read (as in the reference). */
typedef struct trspmm_match_stats {
    int64_t total_rows, total_cols, total_nnz;
    double sparsity;
    int64_t num_row_panels, num_sptc_groups, num_tc_blocks, num_16x16_blocks, block_padding_groups;
    int64_t fine_fallback, coarse_fallback, fake_zeros;
} trspmm_match_stats;
int trspmm_match_2to4(const int32_t* rowptr, const int32_t* col, int64_t rows, int64_t nnz, int window,
                      int dense_threshold, int t_max, int threads, trspmm_match_stats* out);

/* Replaces matching_utils_hash.match_2to4_hash
 * (TR-source/Block/block_matching_hash.cpp:104-441): exact-complement pairs
 * through per-mask chains first, then the sliding-window search with t_max 8;
 * no dense pool (num_tc_blocks is 0, num_sptc_groups is the reference's
 * num_groups).  Host memory, host threads. */
int trspmm_match_2to4_hash(const int32_t* rowptr, const int32_t* col, int64_t rows, int64_t nnz, int window,
                           int threads, trspmm_match_stats* out);

/* Replaces matching_utils_bucket.match_2to4_bucket
 * (TR-source/Block/block_matching_bucket.cpp:72-498): pairs across
 * top-heavy / bottom-heavy / balanced buckets of the columns; no dense pool
 * (num_tc_blocks is 0, num_sptc_groups is the reference's num_groups).  Host
 * memory, host threads (the reference runs it on one thread). */
int trspmm_match_2to4_bucket(const int32_t* rowptr, const int32_t* col, int64_t rows, int64_t nnz, int window,
                             int threads, trspmm_match_stats* out);

/* Replaces matching_utils_online.match_2to4_online
 * (TR-source/Block/block_matching_online.cu:85-649): the match_2to4 heuristic
 * on the device, one workgroup per 16-row window, columns taken in the order
 * of a 2048-slot table (slot = column mod 2048, linear probing; where columns
 * share a slot, the order sequential insertion by ascending column gives),
 * 32 - nnz padding zeros per group.  rowptr [rows+1] and col [nnz] are DEVICE
 * pointers; counters is a device buffer of TRSPMM_ONLINE_COUNTERS unsigned
 * 64-bit values, zeroed and then filled on `stream` (asynchronous):
 *   [0] num_sptc_groups   [1] num_tc_blocks     [2] dense_to_tc_cols
 *   [3] dense_to_sptc_cols [4] fine_fb          [5] coarse_fb_cols
 *   [6] num_16x16_blocks  [7] block_padding_groups [8] fake_zeros
 *   [9] windows with more than 2048 distinct columns (each counted above
 *       with its 2048 smallest column ids, the columns sequential insertion
 *       in ascending order keeps; the reference drops the columns that find
 *       no slot, which ones depending on scheduling)
 *   [10] windows holding a negative column index (not counted above)
 * Row offsets are clamped to [0, nnz], so a malformed rowptr cannot read
 * outside col.  Takes 4 * ceil(rows / 16) + 256 bytes of stream-ordered
 * scratch (hipMallocAsync / hipFreeAsync on `stream`) for the window order. */
#define TRSPMM_ONLINE_COUNTERS 11
int trspmm_match_2to4_online(const int32_t* rowptr, const int32_t* col, int64_t rows, int64_t nnz,
                             int dense_threshold, int t_max, unsigned long long* counters, void* stream);

/* Allocation-free form of trspmm_match_2to4_online: same arguments, same
 * counters, no device allocation and no allocating library call.  It has no
 * counterpart in the reference; it is for callers that cannot allocate
 * (code under HIP graph capture, tight serving loops, harnesses that forbid
 * hipMalloc).  Its scratch (the window order) is a fixed 8 MiB array of
 * device memory per GPU, held until the process ends, so it accepts at most
 * TRSPMM_ONLINE_NOALLOC_MAX_ROWS rows (more return TRSPMM_EINVAL).  Every call
 * on a GPU shares that array: calls that overlap on one GPU (concurrent
 * streams) race and must not be issued. */
#define TRSPMM_ONLINE_NOALLOC_MAX_ROWS (((1LL << 21) - 64) * 16)
int trspmm_match_2to4_online_noalloc(const int32_t* rowptr, const int32_t* col, int64_t rows, int64_t nnz,
                                     int dense_threshold, int t_max, unsigned long long* counters,
                                     void* stream);

/* ------------------------------------------------------------------ TCA row reordering
 * Replaces the row reordering of the reference's caller
 * (spmm/tr_csr_fp16/tr_tca_reorder.py:44-357, apply_tca_reorder): MinHash-LSH
 * candidate pairs, exact Jaccard, two-level greedy clustering (cluster caps
 * `thres`, then `cluster_thres` over the clusters).  Not bit-reproducible
 * with the reference (which hashes through datasketch's SHA-1 MinHash).
 *
 * trspmm_minhash: MinHash signatures on the device, sig[row * num_perm + k]
 * (0xffffffff for an empty row); rowptr / col / sig are device pointers.
 * trspmm_minhash_coefficients: the 2 * num_perm hash coefficients (a_k, b_k)
 * the signatures use (host), so a checker can recompute them.
 * trspmm_tca_order: host arrays; order[new_row] = old row.  `sig` (host,
 * rows * num_perm) may be null: the signatures are then computed on the
 * current device.  Synchronous. */
int trspmm_minhash(const int32_t* rowptr, const int32_t* col, int64_t rows, int num_perm, uint64_t seed,
                   uint32_t* sig, void* stream);
int trspmm_minhash_coefficients(int num_perm, uint64_t seed, uint64_t* out);
int trspmm_tca_order(const int32_t* rowptr, const int32_t* col, int64_t rows, int64_t nnz, const uint32_t* sig,
                     int num_perm, int thres, int cluster_thres, double lsh_threshold, uint64_t seed,
                     int32_t* order);

/* ------------------------------------------------------------------ legacy Libra kernel family
 * Replace the host functions of TR-source/SpMM/mGCNkernel.cu that multiply
 * with the original Libra metadata in fp16 (same parameter lists; device
 * pointers; half -> uint16_t, long -> int64_t):
 *   trspmm_libra_fp16_tcu_cuda           <- spmm_forward_fp16_tcu_cuda_kernel          (:1357-1566)
 *   trspmm_libra_fp16_tcu_cuda_seq       <- spmm_forward_fp16_tcu_cuda_kernel_seq      (:1778-1935)
 *   trspmm_libra_fp16_tcu_cuda_hybrid    <- spmm_forward_fp16_tcu_cuda_hybrid_kernel   (:1936-2100)
 *   trspmm_libra_fp16_tcu_cuda_hybrid_seq<- spmm_forward_fp16_tcu_cuda_hybrid_kernel_seq (:2181-2338)
 *   trspmm_libra_fp16_tcu_cuda_cc_only   <- spmm_forward_fp16_tcu_cuda_kernel_cc_only  (:1567-1777)
 *   trspmm_libra_fp16_tcu_nomap          <- spmm_forward_fp16_tcu_cuda_nomap_kernel    (:2101-2180)
 *   trspmm_libra_fp16_cuda_split         <- spmm_forward_fp16_cuda_kernel_split        (:2651-2766)
 *   trspmm_libra_fp16_cuda_split_v1      <- spmm_forward_fp16_cuda_kernel_split_v1     (:2767-2884)
 * Format: TC blocks of 8-row windows (t_window_row[p] = window index, output
 * rows 8 t_window_row[p] + n; block b = 8 columns t_column[8 b + k] and mask
 * t_binary[b], bit 8 n + k = row n, column k; values at t_blockNew_offset[b]
 * in bit order; parts in window order) and residual parts as in the online
 * format.  Each call builds a plan, runs 10 untimed and `epoches` timed
 * launches on the null stream, leaves the product (fp32, FP32 accumulation,
 * rows [0, mOri), row stride dimN) in output_matrix and returns the mean ms
 * per launch, or -1 with a message for trspmm_last_error().  The seq / hybrid
 * variants, which differ in the reference only in how they schedule the same
 * kernels, compute the same product in one launch; cc_only runs every
 * nonzero on the vector ALU (no matrix instructions); nomap is the TC route
 * alone, split / split_v1 the residual lists alone.  t_atomic / c_atomic,
 * partsize and swizzle are accepted and not needed (every output row has one
 * owner here).  B must hold (largest column + 1) rows of dimN values. */
float trspmm_libra_fp16_tcu_cuda(const int32_t* t_row_offset, const int32_t* t_blockNew_offset,
                                 const uint16_t* t_value, const int32_t* t_column, const int32_t* t_window_row,
                                 const int32_t* t_atomic, const int64_t* t_binary, const int32_t* c_row_offset,
                                 const int32_t* c_row, const int32_t* c_atomic, const int32_t* c_column,
                                 const uint16_t* c_value, const int32_t* c_row_offset_short,
                                 const int32_t* c_row_short, const int32_t* c_atomic_short,
                                 const int32_t* c_column_short, const uint16_t* c_value_short,
                                 const uint16_t* rhs_matrix, float* output_matrix, int parts_t, int parts_c,
                                 int partsize_c, int parts_c_short, int dimN, int mOri, int epoches);
float trspmm_libra_fp16_tcu_cuda_seq(const int32_t* t_row_offset, const int32_t* t_blockNew_offset,
                                     const uint16_t* t_value, const int32_t* t_column, const int32_t* t_window_row,
                                     const int32_t* t_atomic, const int64_t* t_binary, const int32_t* c_row_offset,
                                     const int32_t* c_row, const int32_t* c_atomic, const int32_t* c_column,
                                     const uint16_t* c_value, const int32_t* c_row_offset_short,
                                     const int32_t* c_row_short, const int32_t* c_atomic_short,
                                     const int32_t* c_column_short, const uint16_t* c_value_short,
                                     const uint16_t* rhs_matrix, float* output_matrix, int parts_t, int parts_c,
                                     int partsize_c, int parts_c_short, int dimN, int mOri, int epoches);
float trspmm_libra_fp16_tcu_cuda_hybrid(const int32_t* t_row_offset, const int32_t* t_blockNew_offset,
                                        const uint16_t* t_value, const int32_t* t_column, const int32_t* t_window_row,
                                        const int32_t* t_atomic, const int64_t* t_binary, const int32_t* c_row_offset,
                                        const int32_t* c_row, const int32_t* c_atomic, const int32_t* c_column,
                                        const uint16_t* c_value, const int32_t* c_row_offset_short,
                                        const int32_t* c_row_short, const int32_t* c_atomic_short,
                                        const int32_t* c_column_short, const uint16_t* c_value_short,
                                        const uint16_t* rhs_matrix, float* output_matrix, int parts_t, int parts_c,
                                        int partsize_c, int parts_c_short, int dimN, int mOri, int epoches);
float trspmm_libra_fp16_tcu_cuda_hybrid_seq(const int32_t* t_row_offset, const int32_t* t_blockNew_offset,
                                            const uint16_t* t_value, const int32_t* t_column,
                                            const int32_t* t_window_row, const int32_t* t_atomic,
                                            const int64_t* t_binary, const int32_t* c_row_offset, const int32_t* c_row,
                                            const int32_t* c_atomic, const int32_t* c_column, const uint16_t* c_value,
                                            const int32_t* c_row_offset_short, const int32_t* c_row_short,
                                            const int32_t* c_atomic_short, const int32_t* c_column_short,
                                            const uint16_t* c_value_short, const uint16_t* rhs_matrix,
                                            float* output_matrix, int parts_t, int parts_c, int partsize_c,
                                            int parts_c_short, int dimN, int mOri, int epoches);
float trspmm_libra_fp16_tcu_cuda_cc_only(const int32_t* t_row_offset, const int32_t* t_blockNew_offset,
                                         const uint16_t* t_value, const int32_t* t_column,
                                         const int32_t* t_window_row, const int32_t* t_atomic,
                                         const int64_t* t_binary, const int32_t* c_row_offset, const int32_t* c_row,
                                         const int32_t* c_atomic, const int32_t* c_column, const uint16_t* c_value,
                                         const int32_t* c_row_offset_short, const int32_t* c_row_short,
                                         const int32_t* c_atomic_short, const int32_t* c_column_short,
                                         const uint16_t* c_value_short, const uint16_t* rhs_matrix,
                                         float* output_matrix, int parts_t, int parts_c, int partsize_c,
                                         int parts_c_short, int dimN, int mOri, int epoches);
float trspmm_libra_fp16_tcu_nomap(const int32_t* t_row_offset, const int32_t* t_blockNew_offset,
                                  const uint16_t* t_value, const int32_t* t_column, const int32_t* t_window_row,
                                  const int32_t* t_atomic, const int64_t* t_binary, const uint16_t* rhs_matrix,
                                  float* output_matrix, int parts_t, int dimN, int mOri, int epoches);
float trspmm_libra_fp16_cuda_split(const int32_t* c_row_offset, const int32_t* c_row, const int32_t* c_atomic,
                                   const int32_t* c_column, const uint16_t* c_value, const int32_t* c_row_offset_short,
                                   const int32_t* c_row_short, const int32_t* c_atomic_short,
                                   const int32_t* c_column_short, const uint16_t* c_value_short,
                                   const uint16_t* rhs_matrix, float* output_matrix, int partsize, int dimN, int mOri,
                                   int epoches, int parts, int parts_short, int swizzle);
float trspmm_libra_fp16_cuda_split_v1(const int32_t* c_row_offset, const int32_t* c_row, const int32_t* c_atomic,
                                      const int32_t* c_column, const uint16_t* c_value,
                                      const int32_t* c_row_offset_short, const int32_t* c_row_short,
                                      const int32_t* c_atomic_short, const int32_t* c_column_short,
                                      const uint16_t* c_value_short, const uint16_t* rhs_matrix, float* output_matrix,
                                      int partsize, int dimN, int mOri, int epoches, int parts, int parts_short,
                                      int swizzle);

/* The tf32 members (fp32 values and operands; TC blocks of 8 rows x 4 columns,
 * t_column[4 b + k], a 32-bit mask t_binary[b] whose bit 4 n + k is row n,
 * column k, as spmm_forward_tf32_csr_v2_kernel_tcu reads it, :2886-3030):
 *   trspmm_libra_tf32_tcu_cuda      <- spmm_forward_tf32_tcu_cuda_kernel      (:3513-3674)
 *   trspmm_libra_tf32_tcu_cuda_seq  <- spmm_forward_tf32_tcu_cuda_kernel_seq  (:3675-3837)
 *   trspmm_libra_tf32_cuda_split    <- spmm_forward_tf32_cuda_kernel_split    (:3388-3512)
 * gfx950 has no tf32 matrix instruction: each multiply runs as three bf16
 * products, A_hi B_hi + A_hi B_lo + A_lo B_hi (x_hi = bf16(x), x_lo =
 * bf16(x - x_hi)), summed in fp32 -- finer than tf32's 10-bit mantissa.  The
 * returned ms covers the three products and their sum. */
float trspmm_libra_tf32_tcu_cuda(const int32_t* t_row_offset, const int32_t* t_blockNew_offset, const float* t_value,
                                 const int32_t* t_column, const int32_t* t_window_row, const int32_t* t_atomic,
                                 const int32_t* t_binary, const int32_t* c_row_offset, const int32_t* c_row,
                                 const int32_t* c_atomic, const int32_t* c_column, const float* c_value,
                                 const int32_t* c_row_offset_short, const int32_t* c_row_short,
                                 const int32_t* c_atomic_short, const int32_t* c_column_short,
                                 const float* c_value_short, const float* rhs_matrix, float* output_matrix,
                                 int parts_t, int parts_c, int partsize_c, int parts_c_short, int dimN, int mOri,
                                 int epoches);
float trspmm_libra_tf32_tcu_cuda_seq(const int32_t* t_row_offset, const int32_t* t_blockNew_offset,
                                     const float* t_value, const int32_t* t_column, const int32_t* t_window_row,
                                     const int32_t* t_atomic, const int32_t* t_binary, const int32_t* c_row_offset,
                                     const int32_t* c_row, const int32_t* c_atomic, const int32_t* c_column,
                                     const float* c_value, const int32_t* c_row_offset_short,
                                     const int32_t* c_row_short, const int32_t* c_atomic_short,
                                     const int32_t* c_column_short, const float* c_value_short,
                                     const float* rhs_matrix, float* output_matrix, int parts_t, int parts_c,
                                     int partsize_c, int parts_c_short, int dimN, int mOri, int epoches);
float trspmm_libra_tf32_cuda_split(const int32_t* c_row_offset, const int32_t* c_row, const int32_t* c_atomic,
                                   const int32_t* c_column, const float* c_value, const int32_t* c_row_offset_short,
                                   const int32_t* c_row_short, const int32_t* c_atomic_short,
                                   const int32_t* c_column_short, const float* c_value_short, const float* rhs_matrix,
                                   float* output_matrix, int partsize, int dimN, int mOri, int epoches, int parts,
                                   int parts_short, int swizzle);

/* ------------------------------------------------------------------ diagnostics */

/* libstdc++ std::sort order of iota(n) by key descending, computed by the
 * device router's emulation, run on the host (used by tests). */
int trspmm_host_sort_order(const int32_t* key, int32_t n, int32_t* order);

#ifdef __cplusplus
}
#endif
#endif /* TRSPMM_H */
