// extern "C" entry points of libtrspmm.so (declared in include/trspmm.h).
#include <cstdlib>
#include <cstring>
#include <map>
#include <mutex>
#include <new>
#include <vector>

#include "backward.h"
#include "lsort.h"
#include "plan.h"
#include "router.h"

namespace trspmm {

int match_online_device(const int32_t* d_rowptr, const int32_t* d_col, int64_t rows, int64_t nnz,
                        int dense_threshold, int t_max, unsigned long long* d_out, hipStream_t stream);
int match_online_device_noalloc(const int32_t* d_rowptr, const int32_t* d_col, int64_t rows, int64_t nnz,
                                int dense_threshold, int t_max, unsigned long long* d_out, hipStream_t stream);
int minhash_host_coefficients(int P, uint64_t seed, uint64_t* out);
int minhash_device(const int32_t* d_rowptr, const int32_t* d_col, int64_t rows, int P, uint64_t seed,
                   uint32_t* d_sig, hipStream_t s);
int tca_order(const int32_t* rowptr, const int32_t* col, int64_t rows, int64_t nnz, const uint32_t* sig, int P,
              int thres, int cluster_thres, double lsh_threshold, uint64_t seed, int32_t* order);

static thread_local std::string g_error;
void set_error(const std::string& msg) { g_error = msg; }
const char* last_error() { return g_error.c_str(); }

// ---------------------------------------------------------------- block cache
namespace {
struct CacheRec {
    size_t n;
    int dev;
};
struct BlockCache {
    std::mutex mu;
    std::map<void*, CacheRec> live;                  // every block of the cache, handed out or idle
    std::map<int, std::multimap<size_t, void*>> idle;  // per device, by size
    std::map<int, int64_t> idle_bytes;
    int64_t limit = -1;
    int64_t bound() {
        if (limit < 0) {
            const char* e = std::getenv("TRSPMM_CACHE_MB");
            const long long mb = e ? std::atoll(e) : 4096;
            limit = (int64_t)(mb > 0 ? mb : 0) << 20;
        }
        return limit;
    }
};
BlockCache& block_cache() {
    static BlockCache* c = new BlockCache();         // never destroyed: blocks may be put back at exit
    return *c;
}
}  // namespace

void* cache_get(size_t n) {
    n = (n + 255) & ~(size_t)255;
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess) {
        (void)hipGetLastError();
        set_error("cache_get: no current device");
        return nullptr;
    }
    BlockCache& c = block_cache();
    {
        std::lock_guard<std::mutex> g(c.mu);
        auto& m = c.idle[dev];
        auto it = m.lower_bound(n);
        if (it != m.end() && it->first <= 2 * n) {
            void* p = it->second;
            c.idle_bytes[dev] -= (int64_t)it->first;
            m.erase(it);
            return p;
        }
    }
    void* p = nullptr;
    if (hipMalloc(&p, n) != hipSuccess) {
        (void)hipGetLastError();
        // idle blocks of this device may be what stands in the way
        {
            std::lock_guard<std::mutex> g(c.mu);
            for (auto& kv : c.idle[dev]) {
                (void)hipFree(kv.second);
                c.live.erase(kv.second);
            }
            c.idle[dev].clear();
            c.idle_bytes[dev] = 0;
        }
        if (hipMalloc(&p, n) != hipSuccess) {
            (void)hipGetLastError();
            set_error("hipMalloc failed (" + std::to_string(n) + " bytes)");
            return nullptr;
        }
    }
    std::lock_guard<std::mutex> g(c.mu);
    c.live[p] = CacheRec{n, dev};
    return p;
}

void cache_put(void* p) {
    if (!p) return;
    BlockCache& c = block_cache();
    std::unique_lock<std::mutex> g(c.mu);
    auto it = c.live.find(p);
    if (it == c.live.end()) {          // not a cache block
        g.unlock();
        (void)hipFree(p);
        return;
    }
    const CacheRec r = it->second;
    if (c.idle_bytes[r.dev] + (int64_t)r.n > c.bound()) {
        c.live.erase(it);
        g.unlock();
        (void)hipFree(p);
        return;
    }
    c.idle[r.dev].emplace(r.n, p);
    c.idle_bytes[r.dev] += (int64_t)r.n;
}

int64_t cache_trim() {
    BlockCache& c = block_cache();
    std::lock_guard<std::mutex> g(c.mu);
    int64_t freed = 0;
    for (auto& d : c.idle) {
        for (auto& kv : d.second) {
            (void)hipFree(kv.second);
            c.live.erase(kv.second);
            freed += (int64_t)kv.first;
        }
        d.second.clear();
        c.idle_bytes[d.first] = 0;
    }
    return freed;
}

void* StreamArena::get(size_t n) {
    int best = -1;
    for (int i = 0; i < n_idle; ++i)
        if (idle[i].n >= n && idle[i].n <= 2 * n && (best < 0 || idle[i].n < idle[best].n)) best = i;
    if (best >= 0) {
        void* p = idle[best].p;
        idle[best] = idle[--n_idle];
        return p;
    }
    if (n_held == cap_held) {
        const int nc = cap_held ? 2 * cap_held : 64;
        void** h = (void**)std::realloc(held, sizeof(void*) * nc);
        if (!h) {
            set_error("out of host memory");
            return nullptr;
        }
        held = h;
        cap_held = nc;
    }
    void* p = cache_get(n);
    if (p) held[n_held++] = p;
    return p;
}

void StreamArena::put(void* p, size_t n) {
    if (n_idle == cap_idle) {
        const int nc = cap_idle ? 2 * cap_idle : 64;
        Block* b = (Block*)std::realloc(idle, sizeof(Block) * nc);
        if (!b) return;                // the block stays held; it goes back at the end
        idle = b;
        cap_idle = nc;
    }
    idle[n_idle++] = Block{p, n};
}

StreamArena::~StreamArena() {
    (void)hipStreamSynchronize(s);
    for (int i = 0; i < n_held; ++i) cache_put(held[i]);
    std::free(held);
    std::free(idle);
}

}  // namespace trspmm

using namespace trspmm;

#define TRSPMM_NOTNULL(p, name) TRSPMM_CHECK((p) != nullptr, std::string(name) + " must not be null")

extern "C" {

const char* trspmm_last_error(void) { return trspmm::last_error(); }

int trspmm_version(void) { return 2; }

int trspmm_cache_trim(int64_t* freed) {
    const int64_t n = cache_trim();
    if (freed) *freed = n;
    return TRSPMM_OK;
}

int trspmm_device_arch(char* buf, int len) {
    int dev = 0;
    TRSPMM_HIP(hipGetDevice(&dev));
    hipDeviceProp_t prop;
    TRSPMM_HIP(hipGetDeviceProperties(&prop, dev));
    if (buf && len > 0) {
        std::strncpy(buf, prop.gcnArchName, (size_t)len - 1);
        buf[len - 1] = 0;
    }
    return TRSPMM_OK;
}

// ---------------------------------------------------------------- router
int trspmm_route(const int32_t* d_rowptr, const int32_t* d_col, const float* d_val, int64_t rows, int64_t nnz,
                 int64_t ncols, const trspmm_route_cfg* cfg, void* stream, trspmm_meta** out) {
    TRSPMM_NOTNULL(cfg, "cfg");
    TRSPMM_NOTNULL(out, "out");
    TRSPMM_NOTNULL(d_rowptr, "rowptr");
    TRSPMM_CHECK(nnz == 0 || (d_col && d_val), "col and val must not be null when nnz > 0");
    *out = nullptr;
    trspmm_meta* m = new (std::nothrow) trspmm_meta();
    if (!m) {
        set_error("out of host memory");
        return TRSPMM_ENOMEM;
    }
    const int rc = route_device(d_rowptr, d_col, d_val, rows, nnz, ncols, *cfg, (hipStream_t)stream, m->meta);
    if (rc != TRSPMM_OK) {
        delete m;
        return rc;
    }
    *out = m;
    return TRSPMM_OK;
}

int64_t trspmm_meta_count(const trspmm_meta* meta, int index) {
    if (!meta || index < 0 || index >= 29) return -1;
    return meta->meta.count[index];
}

void* trspmm_meta_data(const trspmm_meta* meta, int index) {
    if (!meta || index < 0 || index >= 29) return nullptr;
    return meta->meta.data[index];
}

double trspmm_meta_seconds(const trspmm_meta* meta) { return meta ? meta->meta.seconds : 0.0; }

const int32_t* trspmm_meta_row_order(const trspmm_meta* meta) { return meta ? meta->meta.row_order : nullptr; }

void trspmm_meta_free(trspmm_meta* meta) { delete meta; }

// ---------------------------------------------------------------- plans
int trspmm_plan_create(const trspmm_meta_view* view, int64_t m_out, int64_t ncols, int dtype, void* stream,
                       trspmm_plan** out) {
    return trspmm_plan_create_ex(view, m_out, ncols, dtype, nullptr, stream, out);
}

int trspmm_plan_create_ex(const trspmm_meta_view* view, int64_t m_out, int64_t ncols, int dtype,
                          const trspmm_plan_opts* opts, void* stream, trspmm_plan** out) {
    TRSPMM_NOTNULL(view, "view");
    TRSPMM_NOTNULL(out, "out");
    *out = nullptr;
    trspmm_plan* p = new (std::nothrow) trspmm_plan();
    if (!p) {
        set_error("out of host memory");
        return TRSPMM_ENOMEM;
    }
    if (opts) p->plan.opts = *opts;
    const int rc = build_plan(*view, m_out, ncols, dtype, (hipStream_t)stream, p->plan);
    if (rc != TRSPMM_OK) {
        delete p;
        return rc;
    }
    *out = p;
    return TRSPMM_OK;
}

int trspmm_plan_from_meta(const trspmm_meta* meta, int64_t m_out, int64_t ncols, int dtype, void* stream,
                          trspmm_plan** out) {
    return trspmm_plan_from_meta_ex(meta, m_out, ncols, dtype, nullptr, stream, out);
}

int trspmm_plan_from_meta_ex(const trspmm_meta* meta, int64_t m_out, int64_t ncols, int dtype,
                             const trspmm_plan_opts* opts, void* stream, trspmm_plan** out) {
    TRSPMM_NOTNULL(meta, "meta");
    const Meta& M = meta->meta;
    const trspmm_meta_view v = meta_view(M);
    if (!M.has_edges && !M.row_order) return trspmm_plan_create_ex(&v, m_out, ncols, dtype, opts, stream, out);
    TRSPMM_NOTNULL(out, "out");
    *out = nullptr;
    trspmm_plan* p = new (std::nothrow) trspmm_plan();
    if (!p) {
        set_error("out of host memory");
        return TRSPMM_ENOMEM;
    }
    if (opts) p->plan.opts = *opts;
    const EdgeMaps em = meta_edges(M);
    const int rc = build_plan(v, m_out, ncols, dtype, (hipStream_t)stream, p->plan, M.has_edges ? &em : nullptr,
                              M.row_order, M.rows);
    if (rc != TRSPMM_OK) {
        delete p;
        return rc;
    }
    *out = p;
    return TRSPMM_OK;
}

int trspmm_plan_set_values(trspmm_plan* plan, const float* vals, int64_t n_vals, const int32_t* index, int64_t nnz,
                           void* stream) {
    TRSPMM_NOTNULL(plan, "plan");
    return plan_set_values(plan->plan, vals, n_vals, index, nnz, (hipStream_t)stream);
}

int trspmm_plan_create_packed(const int32_t* s_tile_column, const int32_t* s_packed_a, const int32_t* s_packed_meta,
                              const int32_t* s_window_tile_offset, int64_t n_tiles, int64_t n_windows, int32_t window,
                              int64_t m_out, int64_t ncols, void* stream, trspmm_plan** out) {
    TRSPMM_NOTNULL(out, "out");
    TRSPMM_NOTNULL(s_window_tile_offset, "s_window_tile_offset");
    *out = nullptr;
    trspmm_plan* p = new (std::nothrow) trspmm_plan();
    if (!p) {
        set_error("out of host memory");
        return TRSPMM_ENOMEM;
    }
    const int rc = build_plan_packed(s_tile_column, s_packed_a, s_packed_meta, s_window_tile_offset, n_tiles,
                                     n_windows, window, m_out, ncols, (hipStream_t)stream, p->plan);
    if (rc != TRSPMM_OK) {
        delete p;
        return rc;
    }
    *out = p;
    return TRSPMM_OK;
}

void trspmm_plan_free(trspmm_plan* plan) { delete plan; }

int trspmm_plan_stats(const trspmm_plan* plan, int64_t* c) {
    TRSPMM_NOTNULL(plan, "plan");
    TRSPMM_NOTNULL(c, "counts");
    const PlanDev& d = plan->plan.dev;
    c[0] = d.windows;
    c[1] = d.n_tsub;
    c[2] = d.n_tval;
    c[3] = d.n_tiles;
    c[4] = d.n_ritems;
    c[5] = d.n_rval;
    return TRSPMM_OK;
}

// ---------------------------------------------------------------- multiply
int trspmm_spmm(const trspmm_plan* plan, const void* B, int64_t ldb, int64_t K, float* C, int64_t ldc,
                uint32_t route_mask, void* stream) {
    TRSPMM_NOTNULL(plan, "plan");
    TRSPMM_NOTNULL(B, "B");
    TRSPMM_NOTNULL(C, "C");
    return launch_spmm(plan->plan.dev, B, ldb, K, C, ldc, route_mask, (hipStream_t)stream);
}

int trspmm_spmm_timed(const trspmm_plan* plan, const void* B, int64_t ldb, int64_t K, float* C, int64_t ldc,
                      uint32_t route_mask, int warmup, int iters, void* stream, float* ms_avg) {
    TRSPMM_NOTNULL(ms_avg, "ms_avg");
    TRSPMM_CHECK(iters > 0 && warmup >= 0, "iters must be positive and warmup non-negative");
    hipStream_t s = (hipStream_t)stream;
    int rc;
    for (int i = 0; i < warmup; ++i)
        if ((rc = trspmm_spmm(plan, B, ldb, K, C, ldc, route_mask, stream))) return rc;
    hipEvent_t e0, e1;
    TRSPMM_HIP(hipEventCreate(&e0));
    TRSPMM_HIP(hipEventCreate(&e1));
    TRSPMM_HIP(hipEventRecord(e0, s));
    for (int i = 0; i < iters; ++i) {
        if ((rc = trspmm_spmm(plan, B, ldb, K, C, ldc, route_mask, stream))) {
            (void)hipEventDestroy(e0);
            (void)hipEventDestroy(e1);
            return rc;
        }
    }
    TRSPMM_HIP(hipEventRecord(e1, s));
    TRSPMM_HIP(hipEventSynchronize(e1));
    float ms = 0.f;
    TRSPMM_HIP(hipEventElapsedTime(&ms, e0, e1));
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    *ms_avg = ms / (float)iters;
    return TRSPMM_OK;
}

int trspmm_memcpy(void* dst, const void* src, int64_t bytes, void* stream) {
    TRSPMM_CHECK(bytes >= 0, "bytes must be non-negative");
    if (bytes == 0) return TRSPMM_OK;
    TRSPMM_CHECK(dst && src, "dst and src must not be null");
    TRSPMM_HIP(hipMemcpyAsync(dst, src, (size_t)bytes, hipMemcpyDefault, (hipStream_t)stream));
    TRSPMM_HIP(hipStreamSynchronize((hipStream_t)stream));
    return TRSPMM_OK;
}

// ---------------------------------------------------------------- diagnostics
// ---------------------------------------------------------------- backward
int trspmm_csr_transpose(const int32_t* d_rowptr, const int32_t* d_col, const float* d_val, int64_t rows,
                         int64_t nnz, int64_t ncols, int32_t* t_rowptr, int32_t* t_col, float* t_val,
                         int32_t* perm, void* stream) {
    return csr_transpose(d_rowptr, d_col, d_val, rows, nnz, ncols, t_rowptr, t_col, t_val, perm, (hipStream_t)stream);
}

int trspmm_sddmm(const int32_t* d_rowptr, int64_t rows, const int32_t* d_col, int64_t nnz, int64_t ncols,
                 const void* X, int64_t ldx, const void* Y, int64_t ldy, int64_t K, int dtype, float* dvals,
                 void* stream) {
    return launch_sddmm(d_rowptr, rows, d_col, nnz, ncols, X, ldx, Y, ldy, K, dtype, dvals, (hipStream_t)stream);
}

int trspmm_bwd_plan_create(const int32_t* t_rowptr, const int32_t* t_col, const float* t_val, const int32_t* perm,
                           int64_t rows, int64_t nnz, int64_t ncols, int dtype, void* stream, trspmm_bwd_plan** out) {
    TRSPMM_NOTNULL(out, "out");
    *out = nullptr;
    trspmm_bwd_plan* p = new (std::nothrow) trspmm_bwd_plan();
    if (!p) {
        set_error("out of host memory");
        return TRSPMM_ENOMEM;
    }
    const int rc = bwd_plan_build(t_rowptr, t_col, t_val, perm, rows, nnz, ncols, dtype, (hipStream_t)stream, p->plan);
    if (rc != TRSPMM_OK) {
        delete p;
        return rc;
    }
    *out = p;
    return TRSPMM_OK;
}

void trspmm_bwd_plan_free(trspmm_bwd_plan* plan) { delete plan; }

int trspmm_bwd_plan_set_values(trspmm_bwd_plan* plan, const float* vals, int64_t nnz, void* stream) {
    TRSPMM_NOTNULL(plan, "plan");
    return bwd_plan_set_values(plan->plan, vals, nnz, (hipStream_t)stream);
}

int trspmm_bwd_plan_stats(const trspmm_bwd_plan* plan, int64_t* counts5) {
    TRSPMM_NOTNULL(plan, "plan");
    TRSPMM_NOTNULL(counts5, "counts");
    const BwdPlan& P = plan->plan;
    counts5[0] = P.routed ? 1 : 0;
    counts5[1] = P.routed ? P.at.dev.n_tsub : 0;
    counts5[2] = P.routed ? P.at.dev.n_tval : 0;
    counts5[3] = P.routed ? P.at.dev.n_rval : 0;
    counts5[4] = (P.routed && P.at.dev.row_map) ? 1 : 0;
    return TRSPMM_OK;
}

int trspmm_spmm_t_sddmm(const trspmm_bwd_plan* plan, const void* dY, int64_t lddy, const void* B, int64_t ldb,
                        int64_t K, float* dB, int64_t lddb, float* dvals, void* stream) {
    TRSPMM_NOTNULL(plan, "plan");
    return launch_bwd(plan->plan, dY, lddy, B, ldb, K, dB, lddb, dvals, (hipStream_t)stream);
}

int trspmm_host_sort_order(const int32_t* key, int32_t n, int32_t* order) {
    TRSPMM_CHECK(n >= 0 && n < (1 << 24), "n must be in [0, 2^24)");
    TRSPMM_CHECK(n == 0 || (key && order), "key and order must not be null");
    std::vector<uint32_t> a((size_t)n);
    for (int32_t i = 0; i < n; ++i) {
        TRSPMM_CHECK(key[i] >= 0 && key[i] < 256, "keys must be in [0, 256)");
        a[(size_t)i] = ((uint32_t)key[i] << 24) | (uint32_t)i;
    }
    int32_t stack[3 * 66];
    lsort::sort(a.data(), n, lsort::KeyDesc{}, stack);
    for (int32_t i = 0; i < n; ++i) order[i] = (int32_t)(a[(size_t)i] & 0xffffffu);
    return TRSPMM_OK;
}

// ---------------------------------------------------------------- matching statistics (device)
int trspmm_match_2to4_online(const int32_t* d_rowptr, const int32_t* d_col, int64_t rows, int64_t nnz,
                             int dense_threshold, int t_max, unsigned long long* d_counters, void* stream) {
    TRSPMM_NOTNULL(d_rowptr, "rowptr");
    TRSPMM_NOTNULL(d_counters, "counters");
    TRSPMM_CHECK(nnz == 0 || d_col, "col must not be null when nnz > 0");
    return match_online_device(d_rowptr, d_col, rows, nnz, dense_threshold, t_max, d_counters, (hipStream_t)stream);
}

int trspmm_match_2to4_online_noalloc(const int32_t* d_rowptr, const int32_t* d_col, int64_t rows, int64_t nnz,
                                     int dense_threshold, int t_max, unsigned long long* d_counters, void* stream) {
    TRSPMM_NOTNULL(d_rowptr, "rowptr");
    TRSPMM_NOTNULL(d_counters, "counters");
    TRSPMM_CHECK(nnz == 0 || d_col, "col must not be null when nnz > 0");
    return match_online_device_noalloc(d_rowptr, d_col, rows, nnz, dense_threshold, t_max, d_counters,
                                       (hipStream_t)stream);
}

// ---------------------------------------------------------------- TCA row reordering
int trspmm_minhash(const int32_t* d_rowptr, const int32_t* d_col, int64_t rows, int num_perm, uint64_t seed,
                   uint32_t* d_sig, void* stream) {
    TRSPMM_NOTNULL(d_rowptr, "rowptr");
    TRSPMM_CHECK(rows == 0 || d_sig, "sig must not be null");
    return minhash_device(d_rowptr, d_col, rows, num_perm, seed, d_sig, (hipStream_t)stream);
}

int trspmm_minhash_coefficients(int num_perm, uint64_t seed, uint64_t* out) {
    TRSPMM_NOTNULL(out, "out");
    TRSPMM_CHECK(num_perm >= 1 && num_perm <= 1024, "num_perm must be in [1, 1024]");
    return minhash_host_coefficients(num_perm, seed, out);
}

int trspmm_tca_order(const int32_t* rowptr, const int32_t* col, int64_t rows, int64_t nnz, const uint32_t* sig,
                     int num_perm, int thres, int cluster_thres, double lsh_threshold, uint64_t seed,
                     int32_t* order) {
    TRSPMM_NOTNULL(rowptr, "rowptr");
    TRSPMM_NOTNULL(order, "order");
    TRSPMM_CHECK(nnz == 0 || col, "col must not be null when nnz > 0");
    try {
        return tca_order(rowptr, col, rows, nnz, sig, num_perm, thres, cluster_thres, lsh_threshold, seed, order);
    } catch (const std::bad_alloc&) {
        set_error("tca_order: out of host memory");
        return TRSPMM_ENOMEM;
    }
}

}  // extern "C"
