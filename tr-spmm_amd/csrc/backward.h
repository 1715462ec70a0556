// Transposition and SDDMM for the backward pass (backward.hip).
#pragma once

#include "common.h"

namespace trspmm {

int csr_transpose(const int32_t* rowptr, const int32_t* col, const float* val, int64_t rows, int64_t nnz,
                  int64_t ncols, int32_t* t_rowptr, int32_t* t_col, float* t_val, int32_t* perm, hipStream_t s);

int launch_sddmm(const int32_t* rowptr, int64_t rows, const int32_t* col, int64_t nnz, const void* X, int64_t ldx,
                 const void* Y, int64_t ldy, int64_t K, int dtype, float* out, hipStream_t s);

}  // namespace trspmm
