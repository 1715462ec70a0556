// TEST INFRASTRUCTURE ONLY — the oracle.
//
// CPU restatement of the reference's Block router (`block_sptc_2to4_online`,
// TR-source/Block/block_online.cpp:545-871) and of the semantics the routed
// metadata encodes (TR-source/SpMM/mGCNkernel_online.cu:106-211, :856-867,
// TR-source/SpMM/spmm_utils/compute.h:53-69).  Only tests/, the smoke check
// and bench.py's CPU-baseline leg load this library, always as the checker.
// The product path (tr-spmm_amd/) never links, loads or calls it.
//
// Two routing strategies, selected by `faithful`:
//   faithful=1  follows the reference step by step: per-column 2-bit row
//               vectors, the pairwise compatibility test, the explicit
//               compatibility graph and its max-degree greedy matching
//               (block_online.cpp:83-98, :267-392).  O(n^2) per window.
//   faithful=0  closed form valid for window <= 16, where every row field of a
//               column vector is 0 or 1 so the compatibility test is always
//               true (block_online.cpp:85-88 with :600, :314): units are
//               consecutive pairs of the (count desc, col asc) list and each
//               idx in the sorted order pairs with the lowest-index unmatched
//               unit.  The order itself still comes from std::sort with the
//               reference's comparator, so libstdc++ tie-breaking is identical.
// Two output modes, selected by `exact`:
//   exact=0     compat: 3rd/4th hits of a 2:4 group row are dropped, exactly as
//               the reference does (block_online.cpp:690-694, :727).
//   exact=1     the dropped hits are appended to the row's residual list
//               (CSR order), so every nonzero is represented exactly once.
//               t_* and s_* are identical in both modes.
//
// Differences from the reference, all deliberate: no printf/cout output; the
// environment knobs LIBRA_SKIP_TC / LIBRA_SPTC_UNIT_MIN_NNZ are explicit
// arguments (the Python wrapper reads the environment the same way the
// reference does, block_online.cpp:118-128, :564-567); window must be <= 16
// (the reference shifts a 32-bit/64-bit word by >= its width for larger
// windows, block_online.cpp:600, :482-486, which is undefined behaviour).

#include <algorithm>
#include <chrono>
#include <cstdint>
#include <cstring>
#include <numeric>
#include <string>
#include <unordered_map>
#include <unordered_set>
#include <vector>

namespace {

struct ColumnStats {               // block_online.cpp:20-25
    int count = 0;
    uint32_t rowvec = 0;           // 2 bits per local row
    std::vector<std::pair<int, int>> hits;   // (local_row, edge), first hit per row
};

struct Unit {                      // block_online.cpp:28-34
    int c0 = -1, c1 = -1;
    uint32_t rowvec = 0;
    int nnz = 0;
};

struct Params {
    int part_t, part_c, short_c, density, window;
    float tc_min_util;
    int skip_tc, unit_min_nnz, exact, faithful, sptc_min_nnz;
};

struct ResidualPart {              // block_online.cpp:37-45
    int count, row, atomic;
    std::vector<int> cols;
    std::vector<float> vals;
};

struct Routed {
    // the 29 arrays, in the order block_online.cpp:844-868 returns them
    std::vector<int32_t> t_row_offset{0}, t_block_offset{0}, t_column;
    std::vector<float> t_value;
    std::vector<int32_t> t_window_row, t_atomic;
    std::vector<int64_t> t_binary;
    std::vector<int32_t> c_row_offset{0}, c_row, c_atomic, c_column;
    std::vector<float> c_value;
    std::vector<int32_t> cs_row_offset{0}, cs_row, cs_atomic, cs_column;
    std::vector<float> cs_value;
    std::vector<int32_t> s_column;
    std::vector<uint8_t> s_mask;
    std::vector<float> s_value;
    std::vector<int8_t> s_pos;
    std::vector<int32_t> s_window, s_offsets{0};
    std::vector<int32_t> s_tile_window, s_tile_group, s_tile_column, s_packed_a, s_packed_meta;
    std::vector<int32_t> s_window_tile_offset{0};
    double duration = 0.0;
    std::string error;
};

// ---- 2-bit row-vector compatibility test (block_online.cpp:83-88) ----
inline bool rowvec_compatible(uint32_t a, uint32_t b) {
    const uint32_t both = a | b;
    const uint32_t overflow = (a & b) | ((both << 1) & both);
    return (overflow & 0xaaaaaaaau) == 0;
}

// IEEE binary16 round-to-nearest-even of a binary32, as c10::Half(float)
// (block_online.cpp:178-183).  Pinned in tests against numpy.float16.
uint16_t f32_to_f16_rne(float f) {
    uint32_t x;
    std::memcpy(&x, &f, 4);
    const uint32_t sign = (x >> 16) & 0x8000u;
    const uint32_t absx = x & 0x7fffffffu;
    if (absx >= 0x7f800000u) {                         // inf / nan
        return (uint16_t)(sign | (absx > 0x7f800000u ? 0x7e00u : 0x7c00u));
    }
    if (absx >= 0x477ff000u) return (uint16_t)(sign | 0x7c00u);   // rounds to inf
    const int exp = (int)(absx >> 23) - 127;
    if (exp >= -14) {                                  // normal half
        uint32_t mant = absx & 0x7fffffu;
        uint32_t h = ((uint32_t)(exp + 15) << 10) | (mant >> 13);
        const uint32_t rem = mant & 0x1fffu;
        if (rem > 0x1000u || (rem == 0x1000u && (h & 1u))) ++h;  // carry may bump exponent: fine
        return (uint16_t)(sign | h);
    }
    if (exp < -25) return (uint16_t)sign;              // below half of min subnormal
    // subnormal half: count units of 2^-24, i.e. mant_full * 2^(exp + 1)
    const uint32_t mant_full = (absx & 0x7fffffu) | 0x800000u;
    const int shift = -(exp + 1);                      // 14 .. 24
    const uint32_t q = mant_full >> shift;
    const uint32_t rem = mant_full & ((1u << shift) - 1u);
    const uint32_t halfway = 1u << (shift - 1);
    uint32_t h = q;
    if (rem > halfway || (rem == halfway && (h & 1u))) ++h;
    return (uint16_t)(sign | h);
}

int32_t pack_half2(float lo, float hi) {
    const uint32_t bits = (uint32_t)f32_to_f16_rne(lo) | ((uint32_t)f32_to_f16_rne(hi) << 16);
    int32_t r;
    std::memcpy(&r, &bits, 4);
    return r;
}

// (count desc, column asc): the total order of block_online.cpp:225-229, :281-285
struct ByCountThenCol {
    const std::unordered_map<int, ColumnStats>* stats;
    bool operator()(int a, int b) const {
        const int ca = stats->at(a).count, cb = stats->at(b).count;
        if (ca != cb) return ca > cb;
        return a < b;
    }
};

// block_online.cpp:215-253.  Returns the accepted TC groups (each sorted asc).
std::vector<std::vector<int>> pick_tc_groups(const std::unordered_map<int, ColumnStats>& stats,
                                             const Params& p, std::unordered_set<int>& tc_cols) {
    std::vector<std::vector<int>> groups;
    std::vector<int> cand;
    for (const auto& kv : stats)
        if (kv.second.count >= p.density) cand.push_back(kv.first);
    std::sort(cand.begin(), cand.end(), ByCountThenCol{&stats});
    const int width = 8;                               // hard-coded at block_online.cpp:611
    for (size_t s = 0; s < cand.size(); s += width) {
        const size_t e = std::min(cand.size(), s + width);
        int nnz = 0;
        std::vector<int> g(cand.begin() + s, cand.begin() + e);
        for (int c : g) nnz += stats.at(c).count;
        const float util = (float)nnz / (float)(p.window * width);
        if (util < p.tc_min_util) continue;
        std::sort(g.begin(), g.end());
        for (int c : g) tc_cols.insert(c);
        groups.push_back(std::move(g));
    }
    return groups;
}

// non-TC columns in (count desc, col asc) order
std::vector<int> sorted_rest(const std::unordered_map<int, ColumnStats>& stats,
                             const std::unordered_set<int>& excluded) {
    std::vector<int> rest;
    for (const auto& kv : stats)
        if (!excluded.count(kv.first)) rest.push_back(kv.first);
    std::sort(rest.begin(), rest.end(), ByCountThenCol{&stats});
    return rest;
}

// block_online.cpp:267-328 (general form: first compatible later column)
std::vector<Unit> units_faithful(const std::unordered_map<int, ColumnStats>& stats,
                                 const std::vector<int>& rest) {
    std::vector<Unit> units;
    std::vector<char> taken(rest.size(), 0);
    for (size_t i = 0; i < rest.size(); ++i) {
        if (taken[i]) continue;
        taken[i] = 1;
        const ColumnStats& a = stats.at(rest[i]);
        Unit u;
        u.c0 = rest[i];
        u.rowvec = a.rowvec;
        u.nnz = a.count;
        for (size_t j = i + 1; j < rest.size(); ++j) {
            if (taken[j]) continue;
            const ColumnStats& b = stats.at(rest[j]);
            if (rowvec_compatible(a.rowvec, b.rowvec)) {
                taken[j] = 1;
                u.c1 = rest[j];
                u.rowvec = a.rowvec | b.rowvec;
                u.nnz = a.count + b.count;
                break;
            }
        }
        units.push_back(u);
    }
    return units;
}

// block_online.cpp:332-392 (general form: explicit graph + greedy)
void match_faithful(const std::vector<Unit>& units, std::vector<std::pair<int, int>>& pairs,
                    std::vector<int>& left) {
    const int n = (int)units.size();
    if (n == 0) return;
    std::vector<std::vector<int>> adj(n);
    std::vector<int> deg(n, 0);
    for (int i = 0; i < n; ++i)
        for (int j = i + 1; j < n; ++j)
            if (rowvec_compatible(units[i].rowvec, units[j].rowvec)) {
                adj[i].push_back(j);
                adj[j].push_back(i);
                ++deg[i];
                ++deg[j];
            }
    std::vector<int> order(n);
    std::iota(order.begin(), order.end(), 0);
    std::sort(order.begin(), order.end(), [&](int a, int b) {
        if (deg[a] != deg[b]) return deg[a] > deg[b];
        return units[a].nnz > units[b].nnz;
    });
    std::vector<char> used(n, 0);
    for (int v : order) {
        if (used[v]) continue;
        int best = -1, best_deg = -1;
        for (int w : adj[v]) {
            if (used[w]) continue;
            if (deg[w] > best_deg || (deg[w] == best_deg && units[w].nnz > units[best].nnz)) {
                best = w;
                best_deg = deg[w];
            }
        }
        if (best >= 0) {
            used[v] = used[best] = 1;
            pairs.push_back({v, best});
        }
    }
    for (int i = 0; i < n; ++i)
        if (!used[i]) left.push_back(i);
}

// Closed form for window <= 16 (complete compatibility graph, see header).
void units_and_match_closed(const std::unordered_map<int, ColumnStats>& stats,
                            const std::vector<int>& rest, std::vector<Unit>& units,
                            std::vector<std::pair<int, int>>& pairs, std::vector<int>& left) {
    for (size_t i = 0; i < rest.size(); i += 2) {
        Unit u;
        u.c0 = rest[i];
        u.nnz = stats.at(rest[i]).count;
        if (i + 1 < rest.size()) {
            u.c1 = rest[i + 1];
            u.nnz += stats.at(rest[i + 1]).count;
        }
        units.push_back(u);
    }
    const int n = (int)units.size();
    if (n == 0) return;
    std::vector<int> order(n);
    std::iota(order.begin(), order.end(), 0);
    // the reference comparator with every degree equal to n-1
    std::sort(order.begin(), order.end(), [&](int a, int b) { return units[a].nnz > units[b].nnz; });
    std::vector<char> used(n, 0);
    int low = 0;                                       // lowest unmatched index
    for (int v : order) {
        if (used[v]) continue;
        while (low < n && used[low]) ++low;
        int w = low;
        if (w == v) {
            w = v + 1;
            while (w < n && used[w]) ++w;
        }
        if (w >= n) continue;
        used[v] = used[w] = 1;
        pairs.push_back({v, w});
    }
    for (int i = 0; i < n; ++i)
        if (!used[i]) left.push_back(i);
}

void route(const int32_t* rowptr, const int32_t* col, const float* val, int64_t rows,
           const Params& p, Routed& out) {
    const auto t0 = std::chrono::high_resolution_clock::now();
    const int W = p.window;
    const int64_t nwin = (rows + W - 1) / W;

    std::vector<ResidualPart> long_parts, short_parts;   // serialised in window order
    std::vector<uint8_t>& s_mask = out.s_mask;

    for (int64_t w = 0; w < nwin; ++w) {
        const int64_t r0 = w * W, r1 = std::min<int64_t>(r0 + W, rows);

        // 1. per-window column compaction (block_online.cpp:588-603)
        std::unordered_map<int, ColumnStats> stats;
        for (int64_t r = r0; r < r1; ++r) {
            const int lr = (int)(r - r0);
            std::unordered_set<int> seen;
            for (int e = rowptr[r]; e < rowptr[r + 1]; ++e) {
                const int c = col[e];
                if (!seen.insert(c).second) continue;
                ColumnStats& cs = stats[c];
                cs.count += 1;
                cs.rowvec += 1u << (2 * lr);
                cs.hits.emplace_back(lr, e);
            }
        }

        // 2. TC columns
        std::unordered_set<int> tc_cols;
        std::vector<std::vector<int>> tc_groups;
        if (!p.skip_tc) tc_groups = pick_tc_groups(stats, p, tc_cols);

        // 3-4. units and their matching
        const std::vector<int> rest = sorted_rest(stats, tc_cols);
        std::vector<Unit> units;
        std::vector<std::pair<int, int>> pairs;
        std::vector<int> left;
        if (p.faithful) {
            units = units_faithful(stats, rest);
            match_faithful(units, pairs, left);
        } else {
            units_and_match_closed(stats, rest, units, pairs, left);
        }

        // 5. 2:4 groups (block_online.cpp:629-666)
        std::unordered_set<int> sp_cols;
        const int g_begin = (int)out.s_window.size();
        auto emit_group = [&](std::vector<int> g) {
            std::sort(g.begin(), g.end());
            for (int c : g) sp_cols.insert(c);
            while (g.size() < 4) g.push_back(-1);
            out.s_window.push_back((int32_t)w);
            out.s_column.insert(out.s_column.end(), g.begin(), g.end());
        };
        for (const auto& pr : pairs) {
            const Unit& a = units[pr.first];
            const Unit& b = units[pr.second];
            std::vector<int> g{a.c0};
            if (a.c1 >= 0) g.push_back(a.c1);
            g.push_back(b.c0);
            if (b.c1 >= 0) g.push_back(b.c1);
            emit_group(g);
        }
        for (int i : left) {
            const Unit& u = units[i];
            if (u.nnz < p.unit_min_nnz) continue;
            std::vector<int> g{u.c0};
            if (u.c1 >= 0) g.push_back(u.c1);
            emit_group(g);
        }

        // 6. per-row slots of each group (block_online.cpp:668-703); spills in exact mode.
        //    Tuned mode (sptc_min_nnz > 0): a group with fewer kept slots leaves the
        //    2:4 route and all of its hits go to the residual route.
        std::unordered_set<int64_t> spilled;         // edge ids moved to the residual route
        const int g_end = (int)out.s_window.size();
        std::vector<int32_t> kept_cols;
        int n_kept = 0;
        for (int g = g_begin; g < g_end; ++g) {
            std::vector<uint8_t> gm(W);
            std::vector<float> gv(2 * W);
            std::vector<int8_t> gp(2 * W);
            std::vector<int64_t> hit_edges, over_edges;
            int kept = 0;
            for (int lr = 0; lr < W; ++lr) {
                uint8_t mask = 0;
                int used = 0;
                float v[2] = {0.f, 0.f};
                int8_t ps[2] = {-1, -1};
                for (int q = 0; q < 4; ++q) {
                    const int c = out.s_column[4 * g + q];
                    if (c < 0) continue;
                    const auto& hits = stats.at(c).hits;
                    auto it = std::find_if(hits.begin(), hits.end(),
                                           [&](const std::pair<int, int>& h) { return h.first == lr; });
                    if (it == hits.end()) continue;
                    mask |= (uint8_t)(1u << q);
                    hit_edges.push_back(it->second);
                    if (used < 2) {
                        v[used] = val[it->second];
                        ps[used] = (int8_t)q;
                        ++kept;
                    } else if (p.exact) {
                        over_edges.push_back(it->second);
                    }
                    ++used;
                }
                gm[lr] = mask;
                gv[2 * lr] = v[0];
                gv[2 * lr + 1] = v[1];
                gp[2 * lr] = ps[0];
                gp[2 * lr + 1] = ps[1];
            }
            if (p.sptc_min_nnz > 0 && kept < p.sptc_min_nnz) {
                spilled.insert(hit_edges.begin(), hit_edges.end());
                continue;
            }
            spilled.insert(over_edges.begin(), over_edges.end());
            kept_cols.insert(kept_cols.end(), out.s_column.begin() + 4 * g, out.s_column.begin() + 4 * g + 4);
            ++n_kept;
            s_mask.insert(s_mask.end(), gm.begin(), gm.end());
            out.s_value.insert(out.s_value.end(), gv.begin(), gv.end());
            out.s_pos.insert(out.s_pos.end(), gp.begin(), gp.end());
        }
        out.s_column.resize(4 * (size_t)g_begin);
        out.s_column.insert(out.s_column.end(), kept_cols.begin(), kept_cols.end());
        out.s_window.resize((size_t)g_begin + n_kept);
        out.s_offsets.push_back((int32_t)out.s_window.size());

        // 7. residual funnel (block_online.cpp:705-733)
        bool has_fallback = false;
        for (const auto& kv : stats)
            if (!tc_cols.count(kv.first) && !sp_cols.count(kv.first)) has_fallback = true;

        // TC sub-blocks and parts (block_online.cpp:444-540)
        if (!tc_groups.empty()) {
            const int nsub = (int)tc_groups.size();
            std::unordered_map<int, int> slot;               // column -> sub*8 + position
            for (int s = 0; s < nsub; ++s)
                for (int q = 0; q < (int)tc_groups[s].size(); ++q) slot[tc_groups[s][q]] = s * 8 + q;
            std::vector<uint64_t> lo(nsub, 0), hi(nsub, 0);
            std::vector<std::vector<float>> sv(nsub);
            for (int64_t r = r0; r < r1; ++r) {
                const int lr = (int)(r - r0);
                for (int e = rowptr[r]; e < rowptr[r + 1]; ++e) {
                    auto f = slot.find(col[e]);
                    if (f == slot.end()) continue;
                    const int s = f->second / 8, q = f->second % 8;
                    if (lr >= 8) hi[s] |= 1ull << ((lr - 8) * 8 + q);
                    else lo[s] |= 1ull << (lr * 8 + q);
                    sv[s].push_back(val[e]);
                }
            }
            for (int s = 0; s < nsub; ++s) {
                for (int q = 0; q < 8; ++q)
                    out.t_column.push_back(q < (int)tc_groups[s].size() ? tc_groups[s][q] : -1);
                out.t_binary.push_back((int64_t)lo[s]);
                out.t_binary.push_back((int64_t)hi[s]);
                out.t_value.insert(out.t_value.end(), sv[s].begin(), sv[s].end());
                out.t_block_offset.push_back(out.t_block_offset.back() + (int32_t)sv[s].size());
            }
            auto add_part = [&](int nsubs, int atomic) {
                out.t_row_offset.push_back(out.t_row_offset.back() + nsubs);
                out.t_window_row.push_back((int32_t)(w * 16));     // block_online.cpp:493
                out.t_atomic.push_back(atomic);
            };
            if (nsub <= p.part_t) {
                add_part(nsub, has_fallback ? 1 : 0);
            } else {
                for (int k = 0; k < nsub / p.part_t; ++k) add_part(p.part_t, 1);
                if (nsub % p.part_t) add_part(nsub % p.part_t, 1);
            }
        }

        // residual rows (block_online.cpp:721-733, split :415-439)
        const int base_atomic = tc_cols.empty() ? 0 : 1;
        for (int64_t r = r0; r < r1; ++r) {
            std::vector<int> rc;
            std::vector<float> rv;
            for (int e = rowptr[r]; e < rowptr[r + 1]; ++e) {
                const int c = col[e];
                const bool routed = tc_cols.count(c) || sp_cols.count(c);
                if (routed && !spilled.count(e)) continue;
                rc.push_back(c);
                rv.push_back(val[e]);
            }
            const int n = (int)rc.size();
            if (n == 0) continue;
            auto add = [&](int off, int cnt, int atomic, bool is_short) {
                ResidualPart part{cnt, (int)r, atomic,
                                  std::vector<int>(rc.begin() + off, rc.begin() + off + cnt),
                                  std::vector<float>(rv.begin() + off, rv.begin() + off + cnt)};
                (is_short ? short_parts : long_parts).push_back(std::move(part));
            };
            if (n <= p.short_c) {
                add(0, n, base_atomic, true);
            } else if (n <= p.part_c) {
                add(0, n, base_atomic, false);
            } else {
                int off = 0;
                for (int k = 0; k < n / p.part_c; ++k, off += p.part_c) add(off, p.part_c, 1, false);
                const int rem = n % p.part_c;
                if (rem) add(off, rem, 1, rem <= p.short_c);
            }
        }
    }

    // serialise residual parts (block_online.cpp:757-776): long list then short list
    for (const auto& pt : long_parts) {
        out.c_row_offset.push_back(out.c_row_offset.back() + pt.count);
        out.c_row.push_back(pt.row);
        out.c_atomic.push_back(pt.atomic);
        out.c_column.insert(out.c_column.end(), pt.cols.begin(), pt.cols.end());
        out.c_value.insert(out.c_value.end(), pt.vals.begin(), pt.vals.end());
    }
    for (const auto& pt : short_parts) {
        out.cs_row_offset.push_back(out.cs_row_offset.back() + pt.count);
        out.cs_row.push_back(pt.row);
        out.cs_atomic.push_back(pt.atomic);
        out.cs_column.insert(out.cs_column.end(), pt.cols.begin(), pt.cols.end());
        out.cs_value.insert(out.cs_value.end(), pt.vals.begin(), pt.vals.end());
    }

    // 2:4 tiles of four consecutive groups (block_online.cpp:796-841)
    for (int64_t w = 0; w < nwin; ++w) {
        const int gb = out.s_offsets[w], ge = out.s_offsets[w + 1];
        for (int base = gb; base < ge; base += 4) {
            out.s_tile_window.push_back((int32_t)w);
            out.s_tile_group.push_back(base);
            for (int k = 0; k < 16; ++k) {
                const int g = base + k / 4;
                out.s_tile_column.push_back(g < ge ? out.s_column[4 * g + k % 4] : -1);
            }
            // one row's (v0,v1,p0,p1) for group g, normalised p0 < p1 (block_online.cpp:190-210)
            auto slot_pair = [&](int g, int row, float& v0, float& v1, int& p0, int& p1) {
                v0 = v1 = 0.f;
                p0 = 0;
                p1 = 1;
                if (g >= ge || row >= W) return;
                const size_t b = ((size_t)g * W + row) * 2;
                const int q0 = out.s_pos[b], q1 = out.s_pos[b + 1];
                const float x0 = out.s_value[b], x1 = out.s_value[b + 1];
                if (q0 < 0 && q1 < 0) return;
                if (q0 >= 0 && q1 >= 0) {
                    if (q0 <= q1) { p0 = q0; p1 = q1; v0 = x0; v1 = x1; }
                    else          { p0 = q1; p1 = q0; v0 = x1; v1 = x0; }
                    return;
                }
                const int q = q0 >= 0 ? q0 : q1;
                const float x = q0 >= 0 ? x0 : x1;
                const int dummy = q == 0 ? 1 : 0;
                if (dummy < q) { p0 = dummy; p1 = q; v0 = 0.f; v1 = x; }
                else           { p0 = q; p1 = dummy; v0 = x; v1 = 0.f; }
            };
            for (int rg = 0; rg < 8; ++rg) {
                uint32_t meta = 0;
                int32_t a_lo[4], a_hi[4];
                for (int ch = 0; ch < 4; ++ch) {
                    float v0, v1;
                    int p0, p1;
                    slot_pair(base + ch, rg, v0, v1, p0, p1);
                    a_lo[ch] = pack_half2(v0, v1);
                    meta |= ((uint32_t)(p0 & 3) | ((uint32_t)(p1 & 3) << 2)) << (4 * ch);
                    slot_pair(base + ch, rg + 8, v0, v1, p0, p1);
                    a_hi[ch] = pack_half2(v0, v1);
                    meta |= ((uint32_t)(p0 & 3) | ((uint32_t)(p1 & 3) << 2)) << (16 + 4 * ch);
                }
                int32_t m;
                std::memcpy(&m, &meta, 4);
                out.s_packed_meta.push_back(m);
                for (int ch = 0; ch < 4; ++ch) {
                    out.s_packed_a.push_back(a_lo[ch]);
                    out.s_packed_a.push_back(a_hi[ch]);
                }
            }
        }
        out.s_window_tile_offset.push_back((int32_t)out.s_tile_window.size());
    }
    const auto t1 = std::chrono::high_resolution_clock::now();
    out.duration = std::chrono::duration<double>(t1 - t0).count();
}

}  // namespace

extern "C" {

// Routes a CSR matrix.  Returns an opaque result (free with oracle_free).
// On invalid arguments the result carries an error string (oracle_error).
void* oracle_route(const int32_t* rowptr, const int32_t* col, const float* val, int64_t rows,
                   int part_t, int part_c, int short_c, int density, int window, int wide,
                   float tc_min_util, int skip_tc, int unit_min_nnz, int exact, int faithful,
                   int sptc_min_nnz) {
    Routed* r = new Routed();
    // block_online.cpp:551-553, narrowed to window <= 16 (see header)
    if (window <= 0 || window > 16 || wide <= 0 || wide > 32) {
        r->error = "window must be in [1, 16] and wide in [1, 32]";
        return r;
    }
    if (part_t <= 0 || part_c <= 0) {
        r->error = "partSize_t and partSize_c must be positive";
        return r;
    }
    if (sptc_min_nnz > 0 && !exact) {
        r->error = "sptc_min_nnz (tuned routing) needs exact mode";
        return r;
    }
    Params p{part_t, part_c, short_c, density, window, tc_min_util, skip_tc, unit_min_nnz, exact, faithful,
             sptc_min_nnz};
    route(rowptr, col, val, rows, p, *r);
    return r;
}

const char* oracle_error(void* h) {
    Routed* r = (Routed*)h;
    return r->error.empty() ? nullptr : r->error.c_str();
}

// Array `idx` (0..28, reference return order) -> pointer + element count.
int oracle_get(void* h, int idx, const void** ptr, int64_t* n) {
    Routed* r = (Routed*)h;
#define OR_CASE(i, v) case i: *ptr = r->v.data(); *n = (int64_t)r->v.size(); return 0;
    switch (idx) {
        OR_CASE(0, t_row_offset) OR_CASE(1, t_block_offset) OR_CASE(2, t_column) OR_CASE(3, t_value)
        OR_CASE(4, t_window_row) OR_CASE(5, t_atomic) OR_CASE(6, t_binary)
        OR_CASE(7, c_row_offset) OR_CASE(8, c_row) OR_CASE(9, c_atomic) OR_CASE(10, c_column) OR_CASE(11, c_value)
        OR_CASE(12, cs_row_offset) OR_CASE(13, cs_row) OR_CASE(14, cs_atomic) OR_CASE(15, cs_column) OR_CASE(16, cs_value)
        OR_CASE(17, s_column) OR_CASE(18, s_mask) OR_CASE(19, s_value) OR_CASE(20, s_pos) OR_CASE(21, s_window)
        OR_CASE(22, s_offsets) OR_CASE(23, s_tile_window) OR_CASE(24, s_tile_group) OR_CASE(25, s_tile_column)
        OR_CASE(26, s_packed_a) OR_CASE(27, s_packed_meta) OR_CASE(28, s_window_tile_offset)
        default: return -1;
    }
#undef OR_CASE
}

double oracle_duration(void* h) { return ((Routed*)h)->duration; }

void oracle_free(void* h) { delete (Routed*)h; }

// Decodes routed metadata into COO triplets (row, col, value) following the
// metadata semantics: TC sub-blocks (bit (r%8)*8+q of lo/hi, values in bit
// order, rows from t_window_row), 2:4 slots (rows window*s_window+lr, column
// s_column[g][pos]), residual parts (c_row).  `cap` bounds the output; the
// return value is the number of triplets (call with cap=0 to size).
int64_t oracle_decode(
    const int32_t* t_row_offset, int64_t n_tparts, const int32_t* t_block_offset, const int32_t* t_column,
    const float* t_value, const int32_t* t_window_row, const int64_t* t_binary,
    const int32_t* c_row_offset, const int32_t* c_row, int64_t n_cparts, const int32_t* c_column, const float* c_value,
    const int32_t* cs_row_offset, const int32_t* cs_row, int64_t n_csparts, const int32_t* cs_column, const float* cs_value,
    const int32_t* s_column, const float* s_value, const int8_t* s_pos, const int32_t* s_window, int64_t n_groups,
    int window, int32_t* out_row, int32_t* out_col, float* out_val, int64_t cap) {
    int64_t n = 0;
    auto put = [&](int r, int c, float v) {
        if (n < cap) { out_row[n] = r; out_col[n] = c; out_val[n] = v; }
        ++n;
    };
    for (int64_t p = 0; p < n_tparts; ++p) {
        for (int s = t_row_offset[p]; s < t_row_offset[p + 1]; ++s) {
            int vi = t_block_offset[s];
            for (int half = 0; half < 2; ++half) {
                const uint64_t bits = (uint64_t)t_binary[2 * s + half];
                for (int b = 0; b < 64; ++b) {
                    if (!((bits >> b) & 1ull)) continue;
                    const int r = t_window_row[p] + half * 8 + b / 8;
                    put(r, t_column[8 * s + b % 8], t_value[vi++]);
                }
            }
        }
    }
    for (int64_t g = 0; g < n_groups; ++g)
        for (int lr = 0; lr < window; ++lr)
            for (int k = 0; k < 2; ++k) {
                const int64_t b = (g * window + lr) * 2 + k;
                if (s_pos[b] < 0) continue;
                put((int)(s_window[g] * window + lr), s_column[4 * g + s_pos[b]], s_value[b]);
            }
    for (int64_t p = 0; p < n_cparts; ++p)
        for (int e = c_row_offset[p]; e < c_row_offset[p + 1]; ++e) put(c_row[p], c_column[e], c_value[e]);
    for (int64_t p = 0; p < n_csparts; ++p)
        for (int e = cs_row_offset[p]; e < cs_row_offset[p + 1]; ++e) put(cs_row[p], cs_column[e], cs_value[e]);
    return n;
}

// libstdc++ std::sort of iota(n) by (key desc) — the order the reference
// matching uses (block_online.cpp:356-363 with equal degrees).
void oracle_sort_order(const int32_t* key, int32_t n, int32_t* order) {
    std::vector<int> o(n);
    std::iota(o.begin(), o.end(), 0);
    std::sort(o.begin(), o.end(), [&](int a, int b) { return key[a] > key[b]; });
    std::memcpy(order, o.data(), sizeof(int32_t) * (size_t)n);
}

void oracle_f32_to_f16(const float* in, uint16_t* out, int64_t n) {
    for (int64_t i = 0; i < n; ++i) out[i] = f32_to_f16_rne(in[i]);
}

}  // extern "C"
