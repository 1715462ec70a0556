// TCA row reordering (the reference's spmm/tr_csr_fp16/tr_tca_reorder.py:44-357,
// after DTC-SpMM's TCA_reorder): rows with similar column sets are clustered so
// that 16-row windows hold rows that share columns, which raises the dense
// TC route's share.  The reference builds on datasketch's MinHash-LSH (not
// available here, and not bit-reproducible across hash choices); this is an
// implementation of the same method:
//   1. MinHash signatures of every row's column set on the device, one wave
//      per row: num_perm hashes h_k(c) = ((a_k * x(c) + b_k) mod (2^61 - 1))
//      mod 2^32 with x(c) a 32-bit mix of the column index (datasketch uses the
//      same universal family over a SHA-1 of the column's decimal string);
//   2. LSH: b bands of r signature rows, (b, r) minimising the integrated
//      false-positive + false-negative probability at the threshold with
//      equal weights (datasketch's _optimal_param); rows sharing a band bucket
//      are candidate pairs;
//   3. candidate pairs scored by exact Jaccard similarity and merged greedily
//      from the most similar, union-find with a cluster size cap (`thres`);
//      a pair whose rows were already merged is re-queued on their roots
//      with the Jaccard of the roots' own rows (tr_tca_reorder.py:137-172);
//   4. the same over the row clusters (their column sets merged), cap 128;
//   5. rows in cluster-of-clusters order.
// Pairs of equal similarity are taken in ascending (lower, higher) index
// order (the reference takes them in its heap's order), and an LSH bucket of
// more than 64 rows links each row to its next 16 rather than to all (the
// reference pairs all, quadratic in the bucket).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <queue>
#include <unordered_map>
#include <vector>

#include "common.h"

namespace trspmm {

namespace {

constexpr uint64_t kMersenne61 = (1ull << 61) - 1;
constexpr size_t kBucketAll = 64;     // LSH buckets up to this size contribute all their pairs
constexpr size_t kBucketLink = 16;    // larger ones link each item to this many successors

__host__ __device__ inline uint32_t mix32(uint32_t x) {     // a 32-bit finaliser (murmur3 fmix32)
    x ^= x >> 16;
    x *= 0x85ebca6bu;
    x ^= x >> 13;
    x *= 0xc2b2ae35u;
    x ^= x >> 16;
    return x;
}

// (a * x + b) mod (2^61 - 1), a, b < 2^61, x < 2^32
__host__ __device__ inline uint64_t mod61_affine(uint64_t a, uint64_t b, uint32_t x) {
    // a * x < 2^93: split into (hi, lo) 64-bit halves, add b, fold 2^61 = 1 (mod p)
#if defined(__HIP_DEVICE_COMPILE__)
    const uint64_t lo = a * (uint64_t)x, hi = __umul64hi(a, (uint64_t)x);
#else
    const unsigned __int128 prod = (unsigned __int128)a * x;
    const uint64_t lo = (uint64_t)prod, hi = (uint64_t)(prod >> 64);
#endif
    const uint64_t lo2 = lo + b, hi2 = hi + (lo2 < lo ? 1 : 0);
    uint64_t r = (lo2 & kMersenne61) + ((lo2 >> 61) | (hi2 << 3));
    r = (r & kMersenne61) + (r >> 61);
    return r >= kMersenne61 ? r - kMersenne61 : r;
}

// sig[row * P + k] = min over the row's columns of h_k; 0xffffffff for an empty row
__global__ __launch_bounds__(256) void k_minhash(const int32_t* __restrict__ rowptr, const int32_t* __restrict__ col,
                                                 int64_t rows, const uint64_t* __restrict__ ab, int P,
                                                 uint32_t* __restrict__ sig) {
    const int64_t row = (int64_t)blockIdx.x * 4 + (threadIdx.x >> 6);
    const int lane = threadIdx.x & 63;
    if (row >= rows) return;
    const int32_t b = rowptr[row], e = rowptr[row + 1];
    for (int k0 = 0; k0 < P; k0 += 64) {
        const int k = k0 + lane;
        uint32_t m = 0xffffffffu;
        if (k < P) {
            const uint64_t a = ab[2 * k], c = ab[2 * k + 1];
            for (int32_t i = b; i < e; ++i) {
                const uint32_t h = (uint32_t)mod61_affine(a, c, mix32((uint32_t)col[i]));
                m = h < m ? h : m;
            }
            sig[row * P + k] = m;
        }
    }
}

// datasketch's _optimal_param: integrate the false-positive probability of
// 1 - (1 - s^r)^b over [0, t] and the false-negative one over [t, 1]
double integrate(double lo, double hi, int b, int r, bool fp) {
    const int n = 200;                  // Simpson, even n
    const double h = (hi - lo) / n;
    auto f = [&](double s) {
        const double p = 1.0 - std::pow(1.0 - std::pow(s, r), b);
        return fp ? p : 1.0 - p;
    };
    double acc = f(lo) + f(hi);
    for (int i = 1; i < n; ++i) acc += f(lo + i * h) * (i & 1 ? 4.0 : 2.0);
    return acc * h / 3.0;
}

void optimal_bands(double t, int P, int* bands, int* rows_per_band) {
    double best = 1e300;
    *bands = 1;
    *rows_per_band = P;
    for (int b = 1; b <= P; ++b)
        for (int r = 1; r <= P / b; ++r) {
            const double err = 0.5 * integrate(0.0, t, b, r, true) + 0.5 * integrate(t, 1.0, b, r, false);
            if (err < best) {
                best = err;
                *bands = b;
                *rows_per_band = r;
            }
        }
}

double jaccard_sorted(const int32_t* a, int na, const int32_t* b, int nb) {
    if (na == 0 || nb == 0) return 0.0;
    int i = 0, j = 0, inter = 0;
    while (i < na && j < nb) {
        if (a[i] < b[j]) ++i;
        else if (a[i] > b[j]) ++j;
        else { ++inter; ++i; ++j; }
    }
    const int uni = na + nb - inter;
    return uni > 0 ? (double)inter / uni : 0.0;
}

struct Cand {
    double simi;
    int32_t p1, p2;
};
struct CandLess {   // max-heap on similarity, ties: smaller (p1, p2) first
    bool operator()(const Cand& x, const Cand& y) const {
        if (x.simi != y.simi) return x.simi < y.simi;
        if (x.p1 != y.p1) return x.p1 > y.p1;
        return x.p2 > y.p2;
    }
};

// One clustering level over items with sorted column lists (sets[i], an
// offset/size into `flat`) and MinHash signatures (sig, P per item).
// Returns cluster roots for every item.
std::vector<int32_t> cluster_level(int64_t n, const std::vector<int64_t>& off, const std::vector<int32_t>& flat,
                                   const std::vector<uint32_t>& sig, int P, int bands, int rpb, int cap) {
    auto size_of = [&](int64_t i) { return (int)(off[i + 1] - off[i]); };
    auto jac = [&](int64_t i, int64_t j) {
        return jaccard_sorted(flat.data() + off[i], size_of(i), flat.data() + off[j], size_of(j));
    };
    // LSH buckets per band: (band key) -> items, candidates = items sharing a bucket
    std::vector<std::pair<int32_t, int32_t>> pairs;
    for (int band = 0; band < bands; ++band) {
        std::unordered_map<uint64_t, std::vector<int32_t>> buckets;
        buckets.reserve((size_t)n);
        for (int64_t i = 0; i < n; ++i) {
            if (size_of(i) == 0) continue;
            uint64_t h = 1469598103934665603ull ^ (uint64_t)band;
            for (int q = 0; q < rpb; ++q) {
                h ^= sig[(size_t)i * P + band * rpb + q];
                h *= 1099511628211ull;
            }
            buckets[h].push_back((int32_t)i);
        }
        for (auto& kv : buckets) {
            // every pair of a bucket; a bucket of more than kBucketAll items (rows
            // with near-identical column sets, e.g. the many one-nonzero rows of a
            // power-law graph) links each item to its next kBucketLink only
            const auto& v = kv.second;
            const size_t reach = v.size() > kBucketAll ? kBucketLink : v.size();
            for (size_t x = 0; x < v.size(); ++x)
                for (size_t y = x + 1; y < v.size() && y <= x + reach; ++y) pairs.emplace_back(v[x], v[y]);
        }
    }
    std::sort(pairs.begin(), pairs.end());
    pairs.erase(std::unique(pairs.begin(), pairs.end()), pairs.end());
    std::priority_queue<Cand, std::vector<Cand>, CandLess> que;
    std::unordered_map<uint64_t, char> queued;
    auto key = [&](int64_t a, int64_t b) {
        if (a > b) std::swap(a, b);
        return (uint64_t)a * (uint64_t)n + (uint64_t)b;
    };
    for (const auto& pr : pairs) {
        que.push(Cand{jac(pr.first, pr.second), pr.first, pr.second});
        queued[key(pr.first, pr.second)] = 1;
    }
    std::vector<int32_t> id(n), sz(n, 1);
    std::vector<char> closed(n, 0);
    for (int64_t i = 0; i < n; ++i) id[i] = (int32_t)i;
    auto root = [&](int32_t i) {
        while (i != id[i]) {
            id[i] = id[id[i]];
            i = id[i];
        }
        return i;
    };
    int64_t live = n;
    while (!que.empty() && live > 0) {
        const Cand c = que.top();
        que.pop();
        queued.erase(key(c.p1, c.p2));
        int32_t p1 = c.p1, p2 = c.p2;
        if (p1 == id[p1] && p2 == id[p2]) {
            if (closed[p1] || closed[p2]) continue;
            if (sz[p1] < sz[p2]) std::swap(p1, p2);      // the smaller joins the larger (ties: p2 joins p1)
            id[p2] = p1;
            --live;
            sz[p1] += sz[p2];
            if (sz[p1] >= cap) {
                closed[p1] = 1;
                --live;
            }
        } else {
            const int32_t r1 = root(p1), r2 = root(p2);
            if (closed[r1] || closed[r2]) continue;
            if (r1 != r2 && !queued.count(key(r1, r2))) {
                que.push(Cand{jac(r1, r2), r1, r2});
                queued[key(r1, r2)] = 1;
            }
        }
    }
    std::vector<int32_t> out(n);
    for (int64_t i = 0; i < n; ++i) out[i] = root((int32_t)i);
    return out;
}

int signatures(const int32_t* rowptr_h, const int32_t* col_h, int64_t rows, int64_t nnz, const std::vector<uint64_t>& ab,
               int P, std::vector<uint32_t>& sig) {
    DevBuf d_rp, d_col, d_ab, d_sig;
    int rc;
    if ((rc = d_rp.alloc(4 * (rows + 1))) || (rc = d_col.alloc(4 * (nnz > 0 ? nnz : 1))) ||
        (rc = d_ab.alloc(8 * ab.size())) || (rc = d_sig.alloc(4 * (size_t)rows * P)))
        return rc;
    TRSPMM_HIP(hipMemcpy(d_rp.p, rowptr_h, 4 * (rows + 1), hipMemcpyHostToDevice));
    if (nnz > 0) TRSPMM_HIP(hipMemcpy(d_col.p, col_h, 4 * nnz, hipMemcpyHostToDevice));
    TRSPMM_HIP(hipMemcpy(d_ab.p, ab.data(), 8 * ab.size(), hipMemcpyHostToDevice));
    if (rows > 0) {
        hipLaunchKernelGGL(k_minhash, dim3((unsigned)((rows + 3) / 4)), dim3(256), 0, nullptr, d_rp.as<int32_t>(),
                           d_col.as<int32_t>(), rows, d_ab.as<uint64_t>(), P, d_sig.as<uint32_t>());
        TRSPMM_HIP(hipGetLastError());
    }
    sig.resize((size_t)rows * P);
    TRSPMM_HIP(hipMemcpy(sig.data(), d_sig.p, 4 * (size_t)rows * P, hipMemcpyDeviceToHost));
    return TRSPMM_OK;
}

// the 2 * P coefficients (a in [1, p), b in [0, p)) from a splitmix64 stream
std::vector<uint64_t> coefficients(int P, uint64_t seed) {
    std::vector<uint64_t> ab(2 * (size_t)P);
    uint64_t s = seed;
    auto next = [&]() {
        uint64_t z = (s += 0x9e3779b97f4a7c15ull);
        z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
        z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
        return z ^ (z >> 31);
    };
    for (int k = 0; k < P; ++k) {
        ab[2 * k] = 1 + next() % (kMersenne61 - 1);
        ab[2 * k + 1] = next() % kMersenne61;
    }
    return ab;
}

}  // namespace

int minhash_host_coefficients(int P, uint64_t seed, uint64_t* out) {
    const auto ab = coefficients(P, seed);
    std::copy(ab.begin(), ab.end(), out);
    return TRSPMM_OK;
}

int minhash_device(const int32_t* d_rowptr, const int32_t* d_col, int64_t rows, int P, uint64_t seed,
                   uint32_t* d_sig, hipStream_t s) {
    TRSPMM_CHECK(P >= 1 && P <= 1024, "num_perm must be in [1, 1024]");
    const auto ab = coefficients(P, seed);
    DevBuf d_ab;
    int rc;
    if ((rc = d_ab.alloc(8 * ab.size()))) return rc;
    TRSPMM_HIP(hipMemcpyAsync(d_ab.p, ab.data(), 8 * ab.size(), hipMemcpyHostToDevice, s));
    if (rows > 0) {
        hipLaunchKernelGGL(k_minhash, dim3((unsigned)((rows + 3) / 4)), dim3(256), 0, s, d_rowptr, d_col, rows,
                           d_ab.as<uint64_t>(), P, d_sig);
        TRSPMM_HIP(hipGetLastError());
    }
    TRSPMM_HIP(hipStreamSynchronize(s));      // the coefficients are released on return
    return TRSPMM_OK;
}

int tca_order(const int32_t* rowptr, const int32_t* col, int64_t rows, int64_t nnz, const uint32_t* sig_in, int P,
              int thres, int cluster_thres, double lsh_threshold, uint64_t seed, int32_t* order) {
    TRSPMM_CHECK(rows >= 0 && nnz >= 0, "invalid sizes");
    TRSPMM_CHECK(P >= 1 && P <= 1024, "num_perm must be in [1, 1024]");
    TRSPMM_CHECK(thres >= 1 && cluster_thres >= 1, "cluster caps must be >= 1");
    TRSPMM_CHECK(lsh_threshold > 0.0 && lsh_threshold < 1.0, "lsh_threshold must be in (0, 1)");
    TRSPMM_CHECK(rows == 0 || (rowptr[0] == 0 && rowptr[rows] == nnz), "rowptr must run from 0 to nnz");
    for (int64_t r = 0; r < rows; ++r) TRSPMM_CHECK(rowptr[r] <= rowptr[r + 1], "rowptr must be non-decreasing");
    int bands = 0, rpb = 0;
    optimal_bands(lsh_threshold, P, &bands, &rpb);
    // row level
    std::vector<uint32_t> sig;
    if (sig_in) sig.assign(sig_in, sig_in + (size_t)rows * P);
    else if (int rc = signatures(rowptr, col, rows, nnz, coefficients(P, seed), P, sig)) return rc;
    std::vector<int64_t> off(rows + 1);
    std::vector<int32_t> flat(col, col + nnz);
    for (int64_t r = 0; r <= rows; ++r) off[r] = rowptr[r];
    for (int64_t r = 0; r < rows; ++r) std::sort(flat.begin() + off[r], flat.begin() + off[r + 1]);
    const std::vector<int32_t> root1 = cluster_level(rows, off, flat, sig, P, bands, rpb, thres);
    // clusters in order of first appearance; members ascending
    std::vector<int32_t> ckey;                       // cluster index -> root
    std::unordered_map<int32_t, int32_t> cindex;     // root -> cluster index
    std::vector<std::vector<int32_t>> members;
    for (int64_t r = 0; r < rows; ++r) {
        auto it = cindex.find(root1[r]);
        if (it == cindex.end()) {
            it = cindex.emplace(root1[r], (int32_t)ckey.size()).first;
            ckey.push_back(root1[r]);
            members.emplace_back();
        }
        members[it->second].push_back((int32_t)r);
    }
    const int64_t nc = (int64_t)ckey.size();
    // cluster level: merged column sets, their signatures = elementwise min of the rows'
    std::vector<int64_t> coff(nc + 1, 0);
    std::vector<int32_t> cflat;
    std::vector<uint32_t> csig((size_t)nc * P, 0xffffffffu);
    for (int64_t c = 0; c < nc; ++c) {
        std::vector<int32_t> u;
        for (int32_t r : members[c]) {
            u.insert(u.end(), flat.begin() + off[r], flat.begin() + off[r + 1]);
            for (int k = 0; k < P; ++k) csig[(size_t)c * P + k] = std::min(csig[(size_t)c * P + k], sig[(size_t)r * P + k]);
        }
        std::sort(u.begin(), u.end());
        u.erase(std::unique(u.begin(), u.end()), u.end());
        cflat.insert(cflat.end(), u.begin(), u.end());
        coff[c + 1] = (int64_t)cflat.size();
    }
    const std::vector<int32_t> root2 = cluster_level(nc, coff, cflat, csig, P, bands, rpb, cluster_thres);
    std::unordered_map<int32_t, int32_t> gindex;
    std::vector<std::vector<int32_t>> groups;
    for (int64_t c = 0; c < nc; ++c) {
        auto it = gindex.find(root2[c]);
        if (it == gindex.end()) {
            it = gindex.emplace(root2[c], (int32_t)groups.size()).first;
            groups.emplace_back();
        }
        groups[it->second].push_back((int32_t)c);
    }
    int64_t at = 0;
    for (const auto& g : groups)
        for (int32_t c : g)
            for (int32_t r : members[c]) order[at++] = r;
    TRSPMM_CHECK(at == rows, "internal: the order does not cover every row");
    return TRSPMM_OK;
}

}  // namespace trspmm
