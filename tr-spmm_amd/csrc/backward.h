// Transposition and SDDMM for the backward pass (backward.hip).
#pragma once

#include "common.h"
#include "plan.h"

namespace trspmm {

int csr_transpose(const int32_t* rowptr, const int32_t* col, const float* val, int64_t rows, int64_t nnz,
                  int64_t ncols, int32_t* t_rowptr, int32_t* t_col, float* t_val, int32_t* perm, hipStream_t s);

// fused A^T backward (dB = A^T dY and dvals = SDDMM(dY, B) in one pass over A^T)
struct BwdPlan {
    DevBuf rowptr, col, val, invperm, dvals_t, items, zero_rows;
    int64_t rows = 0, nnz = 0, ncols = 0, n_items = 0, n_zero = 0;
    int dtype = TRSPMM_F16;
    // routed path (the default): A^T routed with every nonzero a residual record,
    // its multiply computing the value gradient of each record on the side
    // (D), pos[e] = the record of A's edge e, perm = A's edge of each A^T entry
    Plan at;
    DevBuf pos, perm, D;
    bool routed = false;
    // the buffers are blocks of the device block cache: wait for the device
    // (as hipFree would) before they go back
    BwdPlan() = default;
    BwdPlan(const BwdPlan&) = delete;
    BwdPlan& operator=(const BwdPlan&) = delete;
    ~BwdPlan() { (void)hipDeviceSynchronize(); }
};

int bwd_plan_build(const int32_t* t_rowptr, const int32_t* t_col, const float* t_val, const int32_t* perm,
                   int64_t rows, int64_t nnz, int64_t ncols, int dtype, hipStream_t s, BwdPlan& P);
int bwd_plan_set_values(BwdPlan& P, const float* vals, int64_t nnz, hipStream_t s);
int launch_bwd(const BwdPlan& P, const void* dY, int64_t lddy, const void* B, int64_t ldb, int64_t K, float* dB,
               int64_t lddb, float* dvals, hipStream_t s);

int launch_sddmm(const int32_t* rowptr, int64_t rows, const int32_t* col, int64_t nnz, int64_t ncols, const void* X,
                 int64_t ldx, const void* Y, int64_t ldy, int64_t K, int dtype, float* out, hipStream_t s);

}  // namespace trspmm

struct trspmm_bwd_plan {
    trspmm::BwdPlan plan;
};
