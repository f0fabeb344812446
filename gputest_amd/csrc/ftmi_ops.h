// Sparse operators shared by the Krylov and AMG kernels.
#pragma once

#include "ftmi_common.h"

namespace ftmi {

// s + sum_{k<d} coef(k) * f(idx(k)), added in k order. The loop takes kGather entries at a time:
// first their indices and coefficients, then their gathered values, then the sums, so the loads
// of a chunk are in flight together. A loop over one entry at a time waits out two dependent
// memory round trips per entry (index, then value), which makes the short rows of these
// operators latency-bound. The slots of a chunk past d gather f(safe) (safe: any index f may
// read) and are not added, so the arithmetic and its order are those of the one-at-a-time loop.
constexpr int kGather = 4;

template <typename T, class IDX, class COEF, class F>
__device__ inline T gather_sum(T s, int d, int32_t safe, IDX idx, COEF coef, F f) {
    for (int k0 = 0; k0 < d; k0 += kGather) {
        int32_t nb[kGather];
        T a[kGather], v[kGather];
#pragma unroll
        for (int u = 0; u < kGather; ++u) {
            nb[u] = safe;
            a[u] = T(0);
            if (k0 + u < d) {
                nb[u] = idx(k0 + u);
                a[u] = coef(k0 + u);
            }
        }
#pragma unroll
        for (int u = 0; u < kGather; ++u) v[u] = f(nb[u]);
#pragma unroll
        for (int u = 0; u < kGather; ++u)
            if (k0 + u < d) s += a[u] * v[u];
    }
    return s;
}

// Plain CSR (indptr, indices, data): y_i = sum_j a_ij x_j in column order.
template <typename T>
struct CsrOp {
    const int32_t* ptr;
    const int32_t* col;
    const T* val;
    __device__ inline T row(int64_t i, const T* __restrict__ x) const {
        return row_f(i, [&](int32_t j) { return x[j]; });
    }
    // sum_j a_ij f(j), with f giving x_j on the fly
    template <class F>
    __device__ inline T row_f(int64_t i, F f) const {
        const int32_t p = ptr[i];
        return gather_sum(
            T(0), ptr[i + 1] - p, 0, [&](int k) { return col[p + k]; }, [&](int k) { return val[p + k]; }, f);
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
        return row_f(i, [&](int32_t j) { return x[j]; });
    }
    template <class F>
    __device__ inline T row_f(int64_t i, F f) const {
        const int32_t p = ptr[i];
        return gather_sum(
            diag[i] * f(i), ptr[i + 1] - p, (int32_t)i, [&](int k) { return nbr[p + k]; },
            [&](int k) { return off[p + k]; }, f);
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
        return row_f(i, [&](int32_t j) { return x[j]; });
    }
    template <class F>
    __device__ inline T row_f(int64_t i, F f) const {
        const int32_t b = base[i >> 6] + (int32_t)(i & 63);
        return gather_sum(
            diag[i] * f(i), (int)deg[i], (int32_t)i, [&](int k) { return nbr[b + 64 * k]; },
            [&](int k) { return off[b + 64 * k]; }, f);
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
        return row_f(i, [&](int32_t j) { return x[j]; });
    }
    template <class F>
    __device__ inline T row_f(int64_t i, F f) const {
        const T xi = f(i);
        const int32_t p = ptr[i];
        return gather_sum(
            ex[i] * xi, ptr[i + 1] - p, (int32_t)i, [&](int k) { return nbr[p + k]; },
            [&](int k) { return off[p + k]; },
            [&](int32_t j) { return f(j) - xi; });
    }
};

// The difference-form operator with its half-edges in the 64-row slice layout of LduSellOp:
// the same row sums in the same order as LduDiffOp, coalesced loads of nbr/off.
template <typename T>
struct LduSellDiffOp {
    const T* ex;
    const int32_t* base;
    const uint8_t* deg;
    const int32_t* nbr;
    const T* off;
    __device__ inline T row(int64_t i, const T* __restrict__ x) const {
        return row_f(i, [&](int32_t j) { return x[j]; });
    }
    template <class F>
    __device__ inline T row_f(int64_t i, F f) const {
        const T xi = f(i);
        const int32_t b = base[i >> 6] + (int32_t)(i & 63);
        return gather_sum(
            ex[i] * xi, (int)deg[i], (int32_t)i, [&](int k) { return nbr[b + 64 * k]; },
            [&](int k) { return off[b + 64 * k]; }, [&](int32_t j) { return f(j) - xi; });
    }
};

}  // namespace ftmi
