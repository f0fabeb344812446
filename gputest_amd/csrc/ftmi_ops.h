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

// The face operator with its half-edges in 64-row slices: the k-th half-edge of row i sits at
// base[i / 64] + 64 k + i % 64 (k < deg[i]), so the 64 lanes of a wavefront read one coalesced
// run of nbr/off per k instead of 64 short strided rows. Rows keep their half-edge order, so
// y_i is summed exactly as LduOp sums it.
template <typename T>
struct LduSellOp {
    const T* diag;
    const int32_t* base;
    const uint8_t* deg;
    const int32_t* nbr;
    const T* off;
    __device__ inline T row(int64_t i, const T* __restrict__ x) const {
        T s = diag[i] * x[i];
        const int32_t b = base[i >> 6] + (int32_t)(i & 63);
        const int d = deg[i];
        for (int k = 0; k < d; ++k) s += off[b + 64 * k] * x[nbr[b + 64 * k]];
        return s;
    }
    template <class F>
    __device__ inline T row_f(int64_t i, F f) const {
        T s = diag[i] * f(i);
        const int32_t b = base[i >> 6] + (int32_t)(i & 63);
        const int d = deg[i];
        for (int k = 0; k < d; ++k) s += off[b + 64 * k] * f(nbr[b + 64 * k]);
        return s;
    }
};

// The same operator in difference form: y_i = ex_i x_i + sum_k off_k (x_nbr(k) - x_i) with
// ex_i = d_i + sum_k off_k (the row sum). For a nearly singular Laplacian (row sums ~ 0) the
// rounding of FP32 coefficients then multiplies differences of x, so the error of A x scales
// with |A x| instead of |A| |x|.
template <typename T>
struct LduDiffOp {
    const T* ex;
    const int32_t* ptr;
    const int32_t* nbr;
    const T* off;
    __device__ inline T row(int64_t i, const T* __restrict__ x) const {
        const T xi = x[i];
        T s = ex[i] * xi;
        const int32_t e = ptr[i + 1];
        for (int32_t k = ptr[i]; k < e; ++k) s += off[k] * (x[nbr[k]] - xi);
        return s;
    }
    template <class F>
    __device__ inline T row_f(int64_t i, F f) const {
        const T xi = f(i);
        T s = ex[i] * xi;
        const int32_t e = ptr[i + 1];
        for (int32_t k = ptr[i]; k < e; ++k) s += off[k] * (f(nbr[k]) - xi);
        return s;
    }
};

}  // namespace ftmi
