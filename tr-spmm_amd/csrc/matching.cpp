// Statistics of the "sliding window" 2:4 matching heuristic behind the
// reference's setup_24matching op matching_utils.match_2to4
// (TR-source/Block/block_matching_window_para_dr.cpp:84-419).  Host code:
// the op is an offline analysis of a matrix (how many 2:4 groups, dense
// 16x8 blocks and padding zeros the heuristic would produce), not part of the
// routed multiply.  Windows are independent and run on a pool of host
// threads; the per-window counts are summed in window order, so the result
// does not depend on the thread count.
//
// Per window of `window` rows (columns ordered by column index, each with the
// bit mask of the window rows it touches and its nonzero count):
//   1. 1:2 pairs: each unpaired column i takes, among the next 32 columns
//      not yet paired, the first with no common row, else the first with the
//      fewest common rows if that is <= t_max; a column with no partner goes
//      to the window's dense pool (fine fallback);
//   2. 2:4 groups: each unused pair takes, among the next 1024 unused pairs,
//      the first whose rows do not conflict ((both_i & any_j) | (any_i &
//      both_j) == 0); a group adds 2*rows - nnz padding zeros; a pair with
//      no partner sends both columns to the pool (coarse fallback);
//   3. the pool, in order, forms dense blocks of `dense_threshold` columns;
//      the rest go two at a time into padded 2:4 groups (2*rows - nnz zeros);
//   4. groups are packed four to a 16x16 block per window (ceil), the
//      shortfall counted as padding groups.
// Row bits are 16 wide as in the reference: rows 16.. of a wider window set
// bit (r mod 32) truncated to 16 bits (the x86 result of its shift).
#include <algorithm>
#include <atomic>
#include <cstdint>
#include <cstring>
#include <string>
#include <thread>
#include <vector>

#include "../../include/trspmm.h"

namespace trspmm {
void set_error(const std::string& msg);
}

namespace {

struct Col {
    int32_t id;
    uint16_t mask;
    int32_t nnz;
};

struct Pair {
    int32_t a, b;
    uint16_t both, any;
    int32_t nnz;
};

struct WinCounts {
    int64_t groups = 0, tc = 0, blocks16 = 0, padding = 0, fine = 0, coarse = 0, fake = 0;
};

constexpr int kPairSearch = 32;      // columns looked ahead for a 1:2 partner
constexpr int kGroupSearch = 1024;   // pairs looked ahead for a 2:4 partner

inline uint16_t row_bit(int local_row) { return (uint16_t)(1u << (local_row & 31)); }

// distinct columns of rows [r0, r1) with their row masks, by column index
void window_columns(const int32_t* rowptr, const int32_t* col, int64_t r0, int64_t r1,
                    std::vector<std::pair<int32_t, uint16_t>>& scratch, std::vector<Col>& out) {
    scratch.clear();
    for (int64_t r = r0; r < r1; ++r) {
        const uint16_t bit = row_bit((int)(r - r0));
        for (int32_t e = rowptr[r]; e < rowptr[r + 1]; ++e) scratch.emplace_back(col[e], bit);
    }
    std::sort(scratch.begin(), scratch.end(),
              [](const std::pair<int32_t, uint16_t>& x, const std::pair<int32_t, uint16_t>& y) {
                  return x.first < y.first;
              });
    out.clear();
    for (const auto& kv : scratch) {
        if (!out.empty() && out.back().id == kv.first) out.back().mask |= kv.second;
        else out.push_back(Col{kv.first, kv.second, 0});
    }
    for (Col& c : out) c.nnz = __builtin_popcount(c.mask);
}

WinCounts match_window(const std::vector<Col>& cols, int win_rows, int dense_threshold, int t_max) {
    WinCounts wc;
    const int n = (int)cols.size();
    if (n == 0) return wc;
    std::vector<int32_t> pool;
    std::vector<Pair> pairs;
    std::vector<char> taken(n, 0);
    for (int i = 0; i < n; ++i) {
        if (taken[i]) continue;
        int pick = -1, best = t_max + 1;
        const int jend = std::min(n, i + 1 + kPairSearch);
        for (int j = i + 1; j < jend; ++j) {
            if (taken[j]) continue;
            const int cost = __builtin_popcount(cols[i].mask & cols[j].mask);
            if (cost == 0) {
                pick = j;
                break;
            }
            if (cost <= t_max && cost < best) {
                best = cost;
                pick = j;
            }
        }
        taken[i] = 1;
        if (pick < 0) {
            pool.push_back(cols[i].id);
            ++wc.fine;
            continue;
        }
        taken[pick] = 1;
        pairs.push_back(Pair{cols[i].id, cols[pick].id, (uint16_t)(cols[i].mask & cols[pick].mask),
                             (uint16_t)(cols[i].mask | cols[pick].mask), cols[i].nnz + cols[pick].nnz});
    }
    const int np = (int)pairs.size();
    std::vector<char> used(np, 0);
    int64_t groups = 0;
    for (int i = 0; i < np; ++i) {
        if (used[i]) continue;
        used[i] = 1;
        const Pair& p = pairs[i];
        int mate = -1;
        const int jend = std::min(np, i + 1 + kGroupSearch);
        for (int j = i + 1; j < jend; ++j) {
            if (used[j]) continue;
            const uint16_t z = (uint16_t)((p.both & pairs[j].any) | (p.any & pairs[j].both));
            if (z == 0) {
                mate = j;
                break;
            }
        }
        if (mate < 0) {
            pool.push_back(p.a);
            pool.push_back(p.b);
            wc.coarse += 2;
            continue;
        }
        used[mate] = 1;
        ++groups;
        wc.fake += 2 * win_rows - (p.nnz + pairs[mate].nnz);
    }
    // pool: dense blocks, then padded 2:4 groups of at most two real columns
    const int nd = (int)pool.size();
    const int dense = (nd / dense_threshold) * dense_threshold;
    wc.tc = nd / dense_threshold;
    for (int k = dense; k < nd; k += 2) {
        int nnz = 0;
        for (int q = k; q < std::min(nd, k + 2); ++q) {
            const auto it = std::lower_bound(cols.begin(), cols.end(), pool[q],
                                             [](const Col& c, int32_t id) { return c.id < id; });
            if (it != cols.end() && it->id == pool[q]) nnz += it->nnz;
        }
        ++groups;
        wc.fake += 2 * win_rows - nnz;
    }
    wc.groups = groups;
    wc.blocks16 = (groups + 3) / 4;
    wc.padding = wc.blocks16 * 4 - groups;
    return wc;
}

}  // namespace

extern "C" int trspmm_match_2to4(const int32_t* rowptr, const int32_t* col, int64_t rows, int64_t nnz, int window,
                                 int dense_threshold, int t_max, int threads, trspmm_match_stats* out) {
    if (!out || !rowptr || (nnz > 0 && !col)) {
        trspmm::set_error("match_2to4: null pointer");
        return TRSPMM_EINVAL;
    }
    if (window < 1 || window > 64) {
        trspmm::set_error("match_2to4: window_size must be in [1, 64]");
        return TRSPMM_EINVAL;
    }
    if (dense_threshold < 1) {
        trspmm::set_error("match_2to4: dense_threshold must be >= 1");
        return TRSPMM_EINVAL;
    }
    int64_t cols = 0;
    for (int64_t e = 0; e < nnz; ++e) cols = std::max<int64_t>(cols, (int64_t)col[e] + 1);
    if (rows <= 0 || cols <= 0) {
        trspmm::set_error("match_2to4: invalid matrix dimensions");
        return TRSPMM_EINVAL;
    }
    if (rowptr[0] != 0 || rowptr[rows] != nnz) {
        trspmm::set_error("match_2to4: rowptr must run from 0 to nnz");
        return TRSPMM_EINVAL;
    }
    const int64_t W = (rows + window - 1) / window;
    std::vector<WinCounts> per(W);
    std::atomic<int64_t> next{0};
    auto worker = [&]() {
        std::vector<std::pair<int32_t, uint16_t>> scratch;
        std::vector<Col> cs;
        for (;;) {
            const int64_t w0 = next.fetch_add(16);
            if (w0 >= W) break;
            for (int64_t w = w0; w < std::min(W, w0 + 16); ++w) {
                const int64_t r0 = w * window, r1 = std::min(rows, r0 + window);
                window_columns(rowptr, col, r0, r1, scratch, cs);
                per[w] = match_window(cs, (int)(r1 - r0), dense_threshold, t_max);
            }
        }
    };
    int nt = threads > 0 ? threads : (int)std::max(1u, std::min(16u, std::thread::hardware_concurrency()));
    nt = (int)std::min<int64_t>(nt, W);
    std::vector<std::thread> pool;
    for (int t = 1; t < nt; ++t) pool.emplace_back(worker);
    worker();
    for (auto& t : pool) t.join();
    trspmm_match_stats s{};
    s.total_rows = rows;
    s.total_cols = cols;
    s.total_nnz = nnz;
    s.sparsity = 1.0 - (double)nnz / ((double)rows * (double)cols);
    s.num_row_panels = W;
    for (const WinCounts& c : per) {
        s.num_sptc_groups += c.groups;
        s.num_tc_blocks += c.tc;
        s.num_16x16_blocks += c.blocks16;
        s.block_padding_groups += c.padding;
        s.fine_fallback += c.fine;
        s.coarse_fallback += c.coarse;
        s.fake_zeros += c.fake;
    }
    *out = s;
    return TRSPMM_OK;
}
