// Shared definitions for libtrspmm (gfx950, wave64).
#pragma once

#include <hip/hip_runtime.h>
#include <hip/hip_fp16.h>
#include <hip/hip_bf16.h>

#include <cstdint>
#include <cstdio>
#include <string>
#include <utility>

#include "../../include/trspmm.h"

namespace trspmm {

// ---------------------------------------------------------------- errors
void set_error(const std::string& msg);
const char* last_error();

struct Status {
    int code = TRSPMM_OK;
};

#define TRSPMM_HIP(call)                                                                   \
    do {                                                                                   \
        hipError_t e_ = (call);                                                            \
        if (e_ != hipSuccess) {                                                            \
            ::trspmm::set_error(std::string(#call) + ": " + hipGetErrorString(e_));        \
            return TRSPMM_EHIP;                                                            \
        }                                                                                  \
    } while (0)

#define TRSPMM_CHECK(cond, msg)                                                            \
    do {                                                                                   \
        if (!(cond)) {                                                                     \
            ::trspmm::set_error(msg);                                                      \
            return TRSPMM_EINVAL;                                                          \
        }                                                                                  \
    } while (0)

// ---------------------------------------------------------------- device memory
// Block cache, per device.  hipMalloc maps and hipFree unmaps and waits for the
// device (about 0.3 ms a call on MI355X); the router asks for some seventy
// scratch and output buffers per call.  Blocks handed back are kept and handed
// out again (best fit, at most twice the size asked for).  A block handed back
// must be idle: no queued work may still touch it.  The idle bytes kept per
// device are bounded by TRSPMM_CACHE_MB (default 4096); a block that would pass
// the bound is freed at once.
void* cache_get(size_t n);        // nullptr (error set) on failure; n > 0
void cache_put(void* p);          // p from cache_get
int64_t cache_trim();             // frees every idle block; returns the bytes freed

// Scratch of a run of work on one stream: a block released here is handed to
// later requests of the same scope (stream order keeps the two uses apart); the
// destructor waits for the stream, then returns every block to the cache.
struct StreamArena {
    hipStream_t s;
    struct Block {
        void* p;
        size_t n;
    };
    Block* idle = nullptr;          // released in this scope, reusable on s
    int n_idle = 0, cap_idle = 0;
    void** held = nullptr;          // every block taken from the cache
    int n_held = 0, cap_held = 0;
    explicit StreamArena(hipStream_t stream) : s(stream) {}
    StreamArena(const StreamArena&) = delete;
    StreamArena& operator=(const StreamArena&) = delete;
    void* get(size_t n);
    void put(void* p, size_t n);
    ~StreamArena();
};

// Owned device buffer.  From hipMalloc (plans own their memory this way), from a
// StreamArena (scratch), or from the block cache (router results, which wait for
// the device before they give their blocks back).
struct DevBuf {
    void* p = nullptr;
    size_t bytes = 0;
    StreamArena* arena = nullptr;
    bool cached = false;
    int alloc(size_t n) {
        release();
        bytes = n;
        if (n == 0) return TRSPMM_OK;
        if (hipMalloc(&p, n) != hipSuccess) {
            p = nullptr;
            set_error("hipMalloc failed (" + std::to_string(n) + " bytes)");
            return TRSPMM_ENOMEM;
        }
        return TRSPMM_OK;
    }
    int alloc(StreamArena& a, size_t n) {
        release();
        bytes = n;
        if (n == 0) return TRSPMM_OK;
        p = a.get(n);
        if (!p) return TRSPMM_ENOMEM;
        arena = &a;
        return TRSPMM_OK;
    }
    int alloc_cached(size_t n) {
        release();
        bytes = n;
        if (n == 0) return TRSPMM_OK;
        p = cache_get(n);
        if (!p) return TRSPMM_ENOMEM;
        cached = true;
        return TRSPMM_OK;
    }
    void release() {
        if (p) {
            if (arena) arena->put(p, bytes);
            else if (cached) cache_put(p);
            else (void)hipFree(p);
        }
        p = nullptr;
        bytes = 0;
        arena = nullptr;
        cached = false;
    }
    void swap(DevBuf& o) {
        std::swap(p, o.p);
        std::swap(bytes, o.bytes);
        std::swap(arena, o.arena);
        std::swap(cached, o.cached);
    }
    template <typename T> T* as() const { return (T*)p; }
    ~DevBuf() { release(); }
    DevBuf() = default;
    DevBuf(const DevBuf&) = delete;
    DevBuf& operator=(const DevBuf&) = delete;
};

// ---------------------------------------------------------------- vector types
typedef float f32x4 __attribute__((ext_vector_type(4)));
typedef _Float16 f16x8 __attribute__((ext_vector_type(8)));
typedef _Float16 f16x16 __attribute__((ext_vector_type(16)));
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef __bf16 bf16x16 __attribute__((ext_vector_type(16)));
typedef uint32_t u32x4 __attribute__((ext_vector_type(4)));
typedef uint32_t u32x2 __attribute__((ext_vector_type(2)));

// 16-bit storage type tags: the plan and B share one of these.
struct F16 {
    using scalar = _Float16;
    using x8 = f16x8;
    using x16 = f16x16;
    static constexpr int code = TRSPMM_F16;
};
struct BF16 {
    using scalar = __bf16;
    using x8 = bf16x8;
    using x16 = bf16x16;
    static constexpr int code = TRSPMM_BF16;
};

__host__ __device__ inline uint16_t f32_to_bf16_bits(float f) {
    uint32_t x = __builtin_bit_cast(uint32_t, f);
    if ((x & 0x7fffffffu) > 0x7f800000u) return (uint16_t)((x >> 16) | 0x40u);  // quiet nan
    const uint32_t rounding = 0x7fffu + ((x >> 16) & 1u);                       // RNE
    return (uint16_t)((x + rounding) >> 16);
}

// IEEE binary16 round-to-nearest-even of a binary32, bit-exact and independent
// of the hardware rounding/denormal modes (the router's mma.sp packing must
// match c10::Half(float), which is this conversion).
__host__ __device__ inline uint16_t f32_to_f16_rne(float f) {
    const uint32_t x = __builtin_bit_cast(uint32_t, f);
    const uint32_t sign = (x >> 16) & 0x8000u;
    const uint32_t absx = x & 0x7fffffffu;
    if (absx >= 0x7f800000u) return (uint16_t)(sign | (absx > 0x7f800000u ? 0x7e00u : 0x7c00u));
    if (absx >= 0x477ff000u) return (uint16_t)(sign | 0x7c00u);
    const int e = (int)(absx >> 23) - 127;
    if (e >= -14) {
        const uint32_t mant = absx & 0x7fffffu;
        uint32_t h = ((uint32_t)(e + 15) << 10) | (mant >> 13);
        const uint32_t rem = mant & 0x1fffu;
        if (rem > 0x1000u || (rem == 0x1000u && (h & 1u))) ++h;
        return (uint16_t)(sign | h);
    }
    if (e < -25) return (uint16_t)sign;
    const uint32_t full = (absx & 0x7fffffu) | 0x800000u;
    const int sh = -(e + 1);                       // 14 .. 24: units of 2^-24
    uint32_t h = full >> sh;
    const uint32_t rem = full & ((1u << sh) - 1u), half = 1u << (sh - 1);
    if (rem > half || (rem == half && (h & 1u))) ++h;
    return (uint16_t)(sign | h);
}

__device__ inline uint16_t f32_to_f16_bits(float f) {
    return __builtin_bit_cast(uint16_t, (_Float16)f);  // v_cvt_f16_f32: round to nearest even
}

__device__ inline float bits16_to_f32(uint16_t b, int dtype) {
    if (dtype == TRSPMM_BF16) return __builtin_bit_cast(float, (uint32_t)b << 16);
    return (float)__builtin_bit_cast(_Float16, b);
}

// ---------------------------------------------------------------- plan (device layout)
//
// Everything a routed multiply reads, in lane-friendly layouts:
//   windows      W = ceil(rows / 16); window w covers rows [16w, 16w + 16)
//   tc_sub_off   [W+1]  dense 16x8 sub-blocks of window w: [tc_sub_off[w], tc_sub_off[w+1])
//   tc_cols      [S*8]  column of each sub-block slot (-1 = empty)
//   tc_bits      [S*2]  occupancy: bit (r%8)*8+q of word r/8 <=> (row r, slot q) nonzero
//   tc_valoff    [S+1]  first value of each sub-block; values stored in bit order
//   tc_vals      [nnz_tc] 16-bit values (plan dtype)
//   tc_dense     [S*16*8] the same values per (sub-block, row), zeros in empty slots
//   sp_tile_off  [W+1]  2:4 tiles of window w; a tile = 4 groups x 4 columns
//   sp_cols      [T*16] columns (-1 = empty)
//   sp_a         [T*16*4] uint32: (tile, row r, group c) -> two 16-bit values (v0 | v1 << 16)
//   sp_idx       [T*16]  uint16: (tile, row r) -> nibble c = p0 | p1 << 2 (p0 < p1)
//   res_rowptr   [16W+1] residual route as a CSR over the padded rows
//   res_cols     [R] / res_vals [R] (plan dtype)
//   titems       matrix-core work items {window, kind, begin, end}:
//                kind 0 = all work of the window (TC, 2:4 and, with the
//                matrix-core residual engine, residual; plain stores of the
//                whole 16-row tile), 1 / 2 / 3 = a chunk of TC sub-blocks / 2:4
//                tiles / residual entries of a split window (atomic adds)
//   ritems       residual work items {window, kind | lr << 8, begin, end}:
//                kind 0 = whole window, stores its rows; 1 = whole window,
//                adds to rows the tile kernel wrote (read-modify-write);
//                2 = entries [begin, end) of row 16w+lr of a split window (atomic adds)
//   split_windows windows whose rows are zeroed first and only atomically updated
struct PlanDev {
    int64_t windows = 0, m_out = 0, ncols = 0;
    int32_t window = 16;
    int32_t dtype = TRSPMM_F16;
    int64_t n_tsub = 0, n_tval = 0, n_tiles = 0, n_rval = 0;
    const int32_t* tc_sub_off = nullptr;
    const int32_t* tc_cols = nullptr;
    const uint64_t* tc_bits = nullptr;
    const int32_t* tc_valoff = nullptr;
    const uint16_t* tc_vals = nullptr;
    // tc_dense [S*16] rows of 8 16-bit values (16 bytes): row r of sub-block s
    // with zeros in its empty slots, the A-fragment row the tile kernel loads
    const uint32_t* tc_dense = nullptr;
    const int32_t* sp_tile_off = nullptr;
    const int32_t* sp_cols = nullptr;
    const uint32_t* sp_a = nullptr;
    const uint16_t* sp_idx = nullptr;
    const int32_t* res_rowptr = nullptr;
    const int32_t* res_cols = nullptr;
    const uint16_t* res_vals = nullptr;
    // residual entries as 8-byte records col | value << 32 | (row % 16) << 48,
    // the form the matrix-core residual engine reads
    const uint64_t* res_rec = nullptr;
    // residual engine: 1 = matrix cores inside k_tile_routes (one-hot A fragments),
    // 0 = the vector-ALU CSR kernel k_residual
    int32_t res_mfma = 0;
    // features per pass of the tile kernel (64 or 128)
    int32_t slab = 64;
    // work item -> workgroup of the tile kernel: 0 = one workgroup per item,
    // items heaviest first, dealt round robin over the XCDs; 1 = one per item,
    // window order, contiguous item ranges per XCD; 2 = persistent workgroups
    // pulling the heaviest-first items from the queue counters; 3 = as 0 with
    // the items in window order
    int32_t item_order = 0;
    // tile kernel: 0 = deep-pipelined (k_routes), 1 = one sub-step ahead (k_tile_routes)
    int32_t tile_kernel = 0;
    // persistent tile queue: queue[x * kQueueStride] = tickets taken from
    // XCD queue x, queue[8 * kQueueStride] = workgroups that have left; all
    // zero between launches
    int32_t* queue = nullptr;
    const int4* titems = nullptr;
    int64_t n_titems = 0;
    const int4* ritems = nullptr;
    int64_t n_ritems = 0;
    const int32_t* split_windows = nullptr;
    int64_t n_split = 0;
    // plan row 16 w + r -> output row (-1: none), for a reordered routing; null:
    // plan row i is output row i (rows at or beyond m_out dropped)
    const int32_t* row_map = nullptr;
};

// work-item caps: a window is split when it has more TC sub-blocks, 2:4
// tiles or residual entries than these
struct SplitCaps {
    int32_t tc_sub = 64, tiles = 64, res_entries = 2048;
};

// Workgroup index -> window with consecutive windows kept on one XCD.
// gfx950 dispatches workgroup i to XCD (i % 8).  Launch grids are padded to a
// multiple of 8 (xcd_grid); XCD x then owns the contiguous window range
// [x*per, (x+1)*per), so neighbouring windows (which share columns) share an
// L2.  The map is a bijection on [0, grid); callers skip results >= windows.
constexpr int64_t kXcds = 8;
inline int64_t xcd_grid(int64_t windows) { return ((windows + kXcds - 1) / kXcds) * kXcds; }
// persistent tile queue (PlanDev::queue): 8 per-XCD counters and a done
// counter, int32 each, 256 bytes apart
constexpr int kQueueStride = 64;
constexpr int kQueueInts = (int)(kXcds + 1) * kQueueStride;
__device__ inline int64_t xcd_window(int64_t bid, int64_t grid) {
    const int64_t per = grid / kXcds;
    return (bid % kXcds) * per + bid / kXcds;
}

}  // namespace trspmm
