// Statistics of the reference's setup_24matching ops, the 2:4 matching
// heuristics of
//   match_2to4         TR-source/Block/block_matching_window_para_dr.cpp:84-419
//   match_2to4_hash    TR-source/Block/block_matching_hash.cpp:104-441
//   match_2to4_bucket  TR-source/Block/block_matching_bucket.cpp:72-498
// Host code: the ops are an offline analysis of a matrix (how many 2:4
// groups, dense 16x8 blocks and padding zeros a heuristic would produce), not
// part of the routed multiply.  Windows are independent and run on a pool of
// host threads; the per-window counts are summed in window order, so the
// result does not depend on the thread count.
//
// Every variant sees a window of `window` rows as its distinct columns ordered
// by column index, each with the bit mask of the window rows it touches and
// its nonzero count (popcount of the mask).  Row bits are 16 wide as in the
// reference: rows 16.. of a wider window set bit (r mod 32) truncated to 16
// bits (the x86 result of its shift), so they may vanish from the mask.
//
// match_2to4 (sliding window, dense pool):
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
// match_2to4_hash (exact complements first, no pool):
//   1. a chain per row mask lists the window's columns with that mask, the
//      highest column position first; each unpaired column takes the first
//      unpaired column on the chain of its 16-bit complement (a pair whose
//      nonzero count the reference leaves at 0, so its padding is counted as
//      if it were empty);
//   2. the remaining columns pair as in match_2to4 step 1 with t_max 8; a
//      column with no partner becomes a half pair (fine fallback);
//   3. 2:4 groups as in match_2to4 step 2, except that a pair with no
//      partner becomes a group of its own (2*rows - nnz zeros, coarse
//      fallback 2);
//   4. 16x16 blocks as in match_2to4 step 4.
// match_2to4_bucket (top/bottom-half buckets, no pool):
//   1. columns are bucketed by where their nonzeros sit: more in rows 0-7
//      (top), more in rows 8-15 (bottom), or balanced;
//   2. top columns pair with bottom ones, then leftover top and bottom
//      columns with balanced ones, then balanced with balanced: a column
//      looks at up to 16 unpaired candidates of the other bucket, in column
//      order, and takes the first with no common row, else the first with
//      the fewest (<= 8); leftovers become half pairs (fine fallback);
//   3. 2:4 groups as in match_2to4_hash step 3 but looking only 15 pairs
//      ahead, and each group counts 32 - nnz padding zeros whatever the
//      window height;
//   4. 16x16 blocks as in match_2to4 step 4.
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

// two columns (b = -1 for a half pair), the rows both touch and the rows
// either touches, and the nonzero count the padding arithmetic uses
struct Pair {
    int32_t a, b;
    uint16_t both, any;
    int32_t nnz;
};

struct WinCounts {
    int64_t groups = 0, tc = 0, blocks16 = 0, padding = 0, fine = 0, coarse = 0, fake = 0;
};

enum class Variant { Window, Hash, Bucket };

constexpr int kPairSearch = 32;        // columns looked ahead for a 1:2 partner
constexpr int kGroupSearch = 1024;     // pairs looked ahead for a 2:4 partner
constexpr int kBucketSearch = 16;      // unpaired candidates looked at per bucket search
constexpr int kBucketGroupSearch = 15; // pairs looked ahead for a 2:4 partner (bucket)
constexpr int kFixedTmax = 8;          // common-row limit of the hash and bucket variants

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

Pair make_pair(const Col& x, const Col& y) {
    return Pair{x.id, y.id, (uint16_t)(x.mask & y.mask), (uint16_t)(x.mask | y.mask), x.nnz + y.nnz};
}

Pair half_pair(const Col& x) { return Pair{x.id, -1, 0, x.mask, x.nnz}; }

// Among columns i+1 .. i+lookahead not yet taken: the first with no common
// row with column i, else the first with the fewest common rows if that is
// <= t_max; -1 if none.
int partner_ahead(const std::vector<Col>& cols, const std::vector<char>& taken, int i, int t_max) {
    const int n = (int)cols.size();
    int pick = -1, best = t_max + 1;
    const int jend = std::min(n, i + 1 + kPairSearch);
    for (int j = i + 1; j < jend; ++j) {
        if (taken[j]) continue;
        const int cost = __builtin_popcount(cols[i].mask & cols[j].mask);
        if (cost == 0) return j;
        if (cost <= t_max && cost < best) {
            best = cost;
            pick = j;
        }
    }
    return pick;
}

// The same choice over a bucket's candidate list: skips taken candidates and
// `self`, and stops after looking at kBucketSearch of the rest.
int partner_in(const std::vector<Col>& cols, const std::vector<char>& taken, const std::vector<int>& cand, int i,
               int self) {
    int pick = -1, best = kFixedTmax + 1, seen = 0;
    for (int j : cand) {
        if (taken[j] || j == self) continue;
        const int cost = __builtin_popcount(cols[i].mask & cols[j].mask);
        if (cost == 0) return j;
        if (cost <= kFixedTmax && cost < best) {
            best = cost;
            pick = j;
        }
        if (++seen >= kBucketSearch) break;
    }
    return pick;
}

// 2:4 groups: each unused pair takes the first unused pair among the next
// `lookahead` whose rows do not conflict.  on_group(p, q) for a group,
// on_lonely(p) for a pair left without one.
template <class OnGroup, class OnLonely>
void group_pairs(const std::vector<Pair>& pairs, int lookahead, OnGroup on_group, OnLonely on_lonely) {
    const int np = (int)pairs.size();
    std::vector<char> used(np, 0);
    for (int i = 0; i < np; ++i) {
        if (used[i]) continue;
        used[i] = 1;
        const Pair& p = pairs[i];
        int mate = -1;
        const int jend = std::min(np, i + 1 + lookahead);
        for (int j = i + 1; j < jend; ++j) {
            if (used[j]) continue;
            if ((uint16_t)((p.both & pairs[j].any) | (p.any & pairs[j].both)) == 0) {
                mate = j;
                break;
            }
        }
        if (mate < 0) {
            on_lonely(p);
        } else {
            used[mate] = 1;
            on_group(p, pairs[mate]);
        }
    }
}

void pack_blocks(WinCounts& wc) {
    wc.blocks16 = (wc.groups + 3) / 4;
    wc.padding = wc.blocks16 * 4 - wc.groups;
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
        const int pick = partner_ahead(cols, taken, i, t_max);
        taken[i] = 1;
        if (pick < 0) {
            pool.push_back(i);
            ++wc.fine;
            continue;
        }
        taken[pick] = 1;
        pairs.push_back(make_pair(cols[i], cols[pick]));
    }
    // the pool holds column positions; a pair sent back adds both of its columns
    auto position = [&](int32_t id) {
        return (int32_t)(std::lower_bound(cols.begin(), cols.end(), id,
                                          [](const Col& c, int32_t v) { return c.id < v; }) -
                         cols.begin());
    };
    group_pairs(
        pairs, kGroupSearch,
        [&](const Pair& p, const Pair& q) {
            ++wc.groups;
            wc.fake += 2 * win_rows - (p.nnz + q.nnz);
        },
        [&](const Pair& p) {
            pool.push_back(position(p.a));
            pool.push_back(position(p.b));
            wc.coarse += 2;
        });
    // pool: dense blocks, then padded 2:4 groups of at most two real columns
    const int nd = (int)pool.size();
    wc.tc = nd / dense_threshold;
    for (int k = (int)wc.tc * dense_threshold; k < nd; k += 2) {
        int nnz = 0;
        for (int q = k; q < std::min(nd, k + 2); ++q) nnz += cols[pool[q]].nnz;
        ++wc.groups;
        wc.fake += 2 * win_rows - nnz;
    }
    pack_blocks(wc);
    return wc;
}

// `head` has 65536 entries, all -1 on entry and on return
WinCounts match_window_hash(const std::vector<Col>& cols, int win_rows, std::vector<int32_t>& head,
                            std::vector<int32_t>& chain) {
    WinCounts wc;
    const int n = (int)cols.size();
    if (n == 0) return wc;
    chain.assign(n, -1);
    for (int i = 0; i < n; ++i) {          // pushed at the front: a chain runs from the last column back
        chain[i] = head[cols[i].mask];
        head[cols[i].mask] = i;
    }
    std::vector<char> taken(n, 0);
    std::vector<Pair> pairs;
    for (int i = 0; i < n; ++i) {
        if (taken[i]) continue;
        const uint16_t want = (uint16_t)~cols[i].mask;
        for (int j = head[want]; j >= 0; j = chain[j]) {
            if (taken[j] || j == i) continue;
            Pair p = make_pair(cols[i], cols[j]);
            p.nnz = 0;     // the reference does not set it on this path
            pairs.push_back(p);
            taken[i] = taken[j] = 1;
            break;
        }
    }
    for (int i = 0; i < n; ++i) head[cols[i].mask] = -1;
    for (int i = 0; i < n; ++i) {
        if (taken[i]) continue;
        const int pick = partner_ahead(cols, taken, i, kFixedTmax);
        taken[i] = 1;
        if (pick < 0) {
            pairs.push_back(half_pair(cols[i]));
            ++wc.fine;
        } else {
            taken[pick] = 1;
            pairs.push_back(make_pair(cols[i], cols[pick]));
        }
    }
    group_pairs(
        pairs, kGroupSearch,
        [&](const Pair& p, const Pair& q) {
            ++wc.groups;
            wc.fake += 2 * win_rows - (p.nnz + q.nnz);
        },
        [&](const Pair& p) {
            ++wc.groups;
            wc.fake += 2 * win_rows - p.nnz;
            wc.coarse += 2;
        });
    pack_blocks(wc);
    return wc;
}

WinCounts match_window_bucket(const std::vector<Col>& cols) {
    WinCounts wc;
    const int n = (int)cols.size();
    if (n == 0) return wc;
    std::vector<int> top, bottom, balanced;
    for (int i = 0; i < n; ++i) {
        const int t = __builtin_popcount(cols[i].mask & 0x00FFu), b = __builtin_popcount(cols[i].mask & 0xFF00u);
        (t > b ? top : b > t ? bottom : balanced).push_back(i);
    }
    std::vector<char> taken(n, 0);
    std::vector<Pair> pairs;
    auto sweep = [&](const std::vector<int>& from, const std::vector<int>& to, bool same) {
        for (int i : from) {
            if (taken[i]) continue;
            const int j = partner_in(cols, taken, to, i, same ? i : -1);
            if (j < 0) continue;
            pairs.push_back(make_pair(cols[i], cols[j]));
            taken[i] = taken[j] = 1;
        }
    };
    sweep(top, bottom, false);
    sweep(top, balanced, false);
    sweep(bottom, balanced, false);
    sweep(balanced, balanced, true);
    for (int i = 0; i < n; ++i) {
        if (taken[i]) continue;
        pairs.push_back(half_pair(cols[i]));
        ++wc.fine;
    }
    constexpr int kSlots = 32;   // 2 x 16 rows, whatever the window height
    group_pairs(
        pairs, kBucketGroupSearch,
        [&](const Pair& p, const Pair& q) {
            ++wc.groups;
            wc.fake += kSlots - (p.nnz + q.nnz);
        },
        [&](const Pair& p) {
            ++wc.groups;
            wc.fake += kSlots - p.nnz;
            wc.coarse += 2;
        });
    pack_blocks(wc);
    return wc;
}

int validate(const char* op, const int32_t* rowptr, const int32_t* col, int64_t rows, int64_t nnz, int window,
             const trspmm_match_stats* out, int64_t* cols_out) {
    const std::string name(op);
    if (!out || !rowptr || (nnz > 0 && !col)) {
        trspmm::set_error(name + ": null pointer");
        return TRSPMM_EINVAL;
    }
    if (window < 1 || window > 64) {
        trspmm::set_error(name + ": window_size must be in [1, 64]");
        return TRSPMM_EINVAL;
    }
    int64_t cols = 0;
    for (int64_t e = 0; e < nnz; ++e) cols = std::max<int64_t>(cols, (int64_t)col[e] + 1);
    if (rows <= 0 || cols <= 0) {
        trspmm::set_error(name + ": invalid matrix dimensions");
        return TRSPMM_EINVAL;
    }
    if (rowptr[0] != 0 || rowptr[rows] != nnz) {
        trspmm::set_error(name + ": rowptr must run from 0 to nnz");
        return TRSPMM_EINVAL;
    }
    *cols_out = cols;
    return TRSPMM_OK;
}

int run(Variant v, const char* op, const int32_t* rowptr, const int32_t* col, int64_t rows, int64_t nnz, int window,
        int dense_threshold, int t_max, int threads, trspmm_match_stats* out) {
    int64_t cols = 0;
    if (int rc = validate(op, rowptr, col, rows, nnz, window, out, &cols)) return rc;
    const int64_t W = (rows + window - 1) / window;
    std::vector<WinCounts> per(W);
    std::atomic<int64_t> next{0};
    auto worker = [&]() {
        std::vector<std::pair<int32_t, uint16_t>> scratch;
        std::vector<Col> cs;
        std::vector<int32_t> head, chain;
        if (v == Variant::Hash) head.assign(65536, -1);
        for (;;) {
            const int64_t w0 = next.fetch_add(16);
            if (w0 >= W) break;
            for (int64_t w = w0; w < std::min(W, w0 + 16); ++w) {
                const int64_t r0 = w * window, r1 = std::min(rows, r0 + window);
                const int win_rows = (int)(r1 - r0);
                window_columns(rowptr, col, r0, r1, scratch, cs);
                per[w] = v == Variant::Window ? match_window(cs, win_rows, dense_threshold, t_max)
                         : v == Variant::Hash ? match_window_hash(cs, win_rows, head, chain)
                                              : match_window_bucket(cs);
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

}  // namespace

extern "C" int trspmm_match_2to4(const int32_t* rowptr, const int32_t* col, int64_t rows, int64_t nnz, int window,
                                 int dense_threshold, int t_max, int threads, trspmm_match_stats* out) {
    if (dense_threshold < 1) {
        trspmm::set_error("match_2to4: dense_threshold must be >= 1");
        return TRSPMM_EINVAL;
    }
    return run(Variant::Window, "match_2to4", rowptr, col, rows, nnz, window, dense_threshold, t_max, threads, out);
}

extern "C" int trspmm_match_2to4_hash(const int32_t* rowptr, const int32_t* col, int64_t rows, int64_t nnz,
                                      int window, int threads, trspmm_match_stats* out) {
    return run(Variant::Hash, "match_2to4_hash", rowptr, col, rows, nnz, window, 0, kFixedTmax, threads, out);
}

extern "C" int trspmm_match_2to4_bucket(const int32_t* rowptr, const int32_t* col, int64_t rows, int64_t nnz,
                                        int window, int threads, trspmm_match_stats* out) {
    return run(Variant::Bucket, "match_2to4_bucket", rowptr, col, rows, nnz, window, 0, kFixedTmax, threads, out);
}
