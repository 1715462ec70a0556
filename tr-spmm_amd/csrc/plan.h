// Device plan: owned buffers behind a PlanDev view (layout in common.h).
#pragma once

#include "common.h"

namespace trspmm {

// The routed CSR edge behind each value of the router's t_value, c_value,
// c_value_short and s_value arrays (-1 = an empty slot); nnz = the CSR's nonzeros.
struct EdgeMaps {
    const int32_t *t_edge, *c_edge, *cs_edge, *s_edge;
    int64_t nnz;
};

struct Plan {
    PlanDev dev;
    DevBuf tc_sub_off, tc_cols, tc_bits, tc_valoff, tc_vals, tc_dense;
    DevBuf sp_tile_off, sp_cols, sp_a, sp_idx;
    DevBuf res_rowptr, res_cols, res_vals, res_rec;
    DevBuf titems, ritems, split_windows, queue, row_map;
    // edge maps (plans built with EdgeMaps): the CSR edge behind tc_vals[i],
    // the (low, high) halves of sp_a[i], res_vals[i] and res_rec[i]
    // (rec_edge empty when the records are in res_vals order)
    DevBuf tc_edge, sp_edge, res_edge, rec_edge;
    bool has_edges = false;
    int64_t src_nnz = 0;
    SplitCaps caps;
    trspmm_plan_opts opts{};   // requested engine / slab (0 = default)
    int finalize();
    int build_items(hipStream_t s);   // work-item list from the per-window ranges
    // the buffers are blocks of the device block cache: wait for the device
    // (as hipFree would) before they go back
    Plan() = default;
    Plan(const Plan&) = delete;
    Plan& operator=(const Plan&) = delete;
    ~Plan() { (void)hipDeviceSynchronize(); }
};

// row_order (device, order_rows entries, may be null): the routed rows' output
// rows, from a reordered routing (routed row i is output row row_order[i])
int build_plan(const trspmm_meta_view& v, int64_t m_out, int64_t ncols, int dtype, hipStream_t s, Plan& P,
               const EdgeMaps* edges = nullptr, const int32_t* row_order = nullptr, int64_t order_rows = 0);
// new values for a plan with edge maps: value of CSR edge e = vals[index ? index[e] : e]
// (0 outside [0, n_vals)), converted to the plan's type as at build time
int plan_set_values(Plan& P, const float* vals, int64_t n_vals, const int32_t* index, int64_t nnz, hipStream_t s);
int build_plan_packed(const int32_t* s_tile_column, const int32_t* s_packed_a, const int32_t* s_packed_meta,
                      const int32_t* s_window_tile_offset, int64_t T, int64_t W, int window, int64_t m_out,
                      int64_t ncols, hipStream_t s, Plan& P);

// routed multiply (spmm.hip)
int launch_spmm(const PlanDev& p, const void* B, int64_t ldb, int64_t K, float* C, int64_t ldc, uint32_t mask,
                hipStream_t s);
// dB = (the plan's matrix) dY and, for each residual record of the plan's
// window row r and column c, D[record] = <X[r], dY[c]> (the value gradient of
// a plan of A^T, X = the forward's B).  The plan must hold residual records
// only (fused_sddmm_supported).
int fused_sddmm_supported(const PlanDev& p);
int launch_spmm_sddmm(const PlanDev& p, const void* dY, int64_t lddy, int64_t K, float* dB, int64_t lddb,
                      const void* X, int64_t ldx, int64_t x_rows, float* D, hipStream_t s);

}  // namespace trspmm

struct trspmm_plan {
    trspmm::Plan plan;
};
