// Router output: the reference's 29 metadata arrays, owned on the device.
#pragma once

#include "common.h"
#include "plan.h"

namespace trspmm {

// element size of array `i` in the order of block_online.cpp:844-868
inline size_t meta_elem_size(int i) {
    switch (i) {
        case 6: return 8;                       // t_binary int64
        case 18: case 20: return 1;             // s_mask uint8, s_pos int8
        default: return 4;                      // int32 / float32
    }
}

// The arrays are blocks of the device block cache: the destructor waits for the
// device (as hipFree would) before they go back.
struct Meta {
    DevBuf buf[29];
    void* data[29] = {};
    int64_t count[29] = {};
    int window = 16;
    int64_t rows = 0;
    int64_t nnz = 0;
    double seconds = 0.0;
    int alloc(int i, int64_t n) {
        count[i] = n;
        int rc = buf[i].alloc_cached(meta_elem_size(i) * (size_t)(n > 0 ? n : 0));
        data[i] = buf[i].p;
        return rc;
    }
    // optional edge maps (TRSPMM_ROUTE_EDGES): the CSR edge behind each value of
    // t_value [0], c_value [1], c_value_short [2] and s_value [3] (-1 = empty slot)
    DevBuf edge_buf[4];
    void* edge_data[4] = {};
    int64_t edge_count[4] = {};
    bool has_edges = false;
    int edge_alloc(int i, int64_t n) {
        has_edges = true;
        edge_count[i] = n;
        int rc = edge_buf[i].alloc_cached(4 * (size_t)(n > 0 ? n : 0));
        edge_data[i] = edge_buf[i].p;
        return rc;
    }
    // TRSPMM_ROUTE_REORDER, when the reordered routing was kept: the rows in the
    // order they were routed (routed row i is row_order[i] of the input, `rows`
    // entries); null when the rows kept their order
    DevBuf row_order_buf;
    const int32_t* row_order = nullptr;
    void swap(Meta& o) {
        for (int i = 0; i < 29; ++i) {
            buf[i].swap(o.buf[i]);
            std::swap(data[i], o.data[i]);
            std::swap(count[i], o.count[i]);
        }
        for (int i = 0; i < 4; ++i) {
            edge_buf[i].swap(o.edge_buf[i]);
            std::swap(edge_data[i], o.edge_data[i]);
            std::swap(edge_count[i], o.edge_count[i]);
        }
        std::swap(has_edges, o.has_edges);
        std::swap(window, o.window);
        std::swap(rows, o.rows);
        std::swap(nnz, o.nnz);
        std::swap(seconds, o.seconds);
        row_order_buf.swap(o.row_order_buf);
        std::swap(row_order, o.row_order);
    }
    Meta() = default;
    Meta(const Meta&) = delete;
    Meta& operator=(const Meta&) = delete;
    ~Meta() { (void)hipDeviceSynchronize(); }
};

int route_device(const int32_t* rowptr, const int32_t* col, const float* val, int64_t rows, int64_t nnz,
                 int64_t ncols, const trspmm_route_cfg& cfg, hipStream_t s, Meta& M);

// device views of a router result's arrays (the op's inputs, fp32 values) and
// of its edge maps (when it has them)
trspmm_meta_view meta_view(const Meta& M);
EdgeMaps meta_edges(const Meta& M);

}  // namespace trspmm

struct trspmm_meta {
    trspmm::Meta meta;
};
