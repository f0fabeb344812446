// Sparse operators shared by the Krylov and AMG kernels.
#pragma once

#include "ftmi_common.h"

namespace ftmi {

// Plain CSR (indptr, indices, data): y_i = sum_j a_ij x_j in column order.
template <typename T>
struct CsrOp {
    const int32_t* ptr;
    const int32_t* col;
    const T* val;
    __device__ inline T row(int64_t i, const T* __restrict__ x) const {
        T s = T(0);
        const int32_t e = ptr[i + 1];
        for (int32_t j = ptr[i]; j < e; ++j) s += val[j] * x[col[j]];
        return s;
    }
    // sum_j a_ij f(j), with f giving x_j on the fly
    template <class F>
    __device__ inline T row_f(int64_t i, F f) const {
        T s = T(0);
        const int32_t e = ptr[i + 1];
        for (int32_t j = ptr[i]; j < e; ++j) s += val[j] * f(col[j]);
        return s;
    }
};

// Cell-cell operator of a face discretisation: a diagonal and one coefficient per
// (cell, interior face) half-edge, the half-edges of cell i being ptr[i]..ptr[i+1]
// with neighbour cells nbr[]: y_i = d_i x_i + sum_k off_k x_nbr(k).
template <typename T>
struct LduOp {
    const T* diag;
    const int32_t* ptr;
    const int32_t* nbr;
    const T* off;
    __device__ inline T row(int64_t i, const T* __restrict__ x) const {
        T s = diag[i] * x[i];
        const int32_t e = ptr[i + 1];
        for (int32_t k = ptr[i]; k < e; ++k) s += off[k] * x[nbr[k]];
        return s;
    }
    template <class F>
    __device__ inline T row_f(int64_t i, F f) const {
        T s = diag[i] * f(i);
        const int32_t e = ptr[i + 1];
        for (int32_t k = ptr[i]; k < e; ++k) s += off[k] * f(nbr[k]);
        return s;
    }
};

}  // namespace ftmi
