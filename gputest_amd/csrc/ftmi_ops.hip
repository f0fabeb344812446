// Drop-in finite-volume operators: the reference's native C ABI (flowtrace/native/fvm_kernels.h:27-66)
// re-implemented for gfx950 on device pointers. Semantics and summation order per row follow the
// reference's kernels (flowtrace/native/fvm_kernels.cpp:51-144); indices are int32.
#include <cstdio>
#include <cstring>

#include "ftmi_common.h"
#include "../../include/ftmi.h"

using namespace ftmi;

namespace ftmi {
static thread_local char g_err[512] = "";
void set_error(const char* where, hipError_t e) {
    snprintf(g_err, sizeof(g_err), "%s: %s", where, hipGetErrorString(e));
}
void set_error_msg(const char* msg) { snprintf(g_err, sizeof(g_err), "%s", msg); }
static int g_grid_cap = kMaxGrid;
int grid_cap() { return g_grid_cap; }
}  // namespace ftmi

extern "C" {

int ftmi_abi_version(void) { return FTMI_ABI_VERSION; }

const char* ftmi_last_error(void) { return ftmi::g_err; }

int ftmi_set_grid_cap(int cap) {
    const int was = ftmi::g_grid_cap;
    if (cap >= 64 && cap <= 65536) ftmi::g_grid_cap = cap;
    return was;
}

int ftmi_device_count(void) {
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess) return 0;
    return n;
}

}  // extern "C"

namespace {

__global__ void __launch_bounds__(kBlock) k_cast_f32(int64_t n, const double* __restrict__ x, float* __restrict__ y) {
    int64_t lo, hi;
    block_range(n, lo, hi);
    for (int64_t i = lo + threadIdx.x; i < hi; i += kBlock) y[i] = (float)x[i];
}

// soff[pos[k]] = off[k]: half-edge values into the 64-row slice layout
template <typename T>
__global__ void __launch_bounds__(kBlock) k_slice(int64_t n, const int32_t* __restrict__ pos,
                                                  const T* __restrict__ off, T* __restrict__ soff) {
    int64_t lo, hi;
    block_range(n, lo, hi);
    for (int64_t k = lo + threadIdx.x; k < hi; k += kBlock) soff[pos[k]] = off[k];
}

// y[j] = x[idx[j]]
__global__ void __launch_bounds__(kBlock) k_gather_f64(int64_t n, const int64_t* __restrict__ idx,
                                                       const double* __restrict__ x, double* __restrict__ y) {
    int64_t lo, hi;
    block_range(n, lo, hi);
    for (int64_t j = lo + threadIdx.x; j < hi; j += kBlock) y[j] = x[idx[j]];
}

// y[i, :] = sum_j A[i, j] X[j, :]   (k right-hand sides, CSR order as ft_csr_spmm)
template <int K>
__global__ void __launch_bounds__(kBlock) k_csr_spmm(int64_t n, const int32_t* __restrict__ ptr,
                                                     const int32_t* __restrict__ col, const double* __restrict__ val,
                                                     const double* __restrict__ X, double* __restrict__ Y, int kdyn) {
    int64_t lo, hi;
    block_range(n, lo, hi);
    const int k = K > 0 ? K : kdyn;
    for (int64_t i = lo + threadIdx.x; i < hi; i += kBlock) {
        if (K == 1) {
            double s = 0.0;
            for (int32_t j = ptr[i]; j < ptr[i + 1]; ++j) s += val[j] * X[col[j]];
            Y[i] = s;
        } else {
            for (int c = 0; c < k; ++c) Y[i * k + c] = 0.0;
            for (int32_t j = ptr[i]; j < ptr[i + 1]; ++j) {
                const double a = val[j];
                const double* x = X + (int64_t)col[j] * k;
                for (int c = 0; c < k; ++c) Y[i * k + c] += a * x[c];
            }
        }
    }
}

__global__ void __launch_bounds__(kBlock) k_face_interp(int64_t nf, int k, const double* __restrict__ w,
                                                        const int32_t* __restrict__ P, const int32_t* __restrict__ N,
                                                        const double* __restrict__ phi, double* __restrict__ out) {
    int64_t lo, hi;
    block_range(nf * k, lo, hi);
    for (int64_t t = lo + threadIdx.x; t < hi; t += kBlock) {
        const int64_t f = t / k;
        const int c = (int)(t - f * k);
        const double wf = w[f], wn = 1.0 - wf;
        out[t] = wf * phi[(int64_t)P[f] * k + c] + wn * phi[(int64_t)N[f] * k + c];
    }
}

template <int DIM>
__global__ void __launch_bounds__(kBlock) k_ls_gradient(int64_t nc, const int32_t* __restrict__ iptr,
                                                        const int32_t* __restrict__ iidx,
                                                        const int32_t* __restrict__ bptr,
                                                        const int32_t* __restrict__ bidx,
                                                        const int32_t* __restrict__ P, const int32_t* __restrict__ N,
                                                        const double* __restrict__ wl, const double* __restrict__ d,
                                                        const double* __restrict__ wlb, const double* __restrict__ db,
                                                        const double* __restrict__ Ginv,
                                                        const double* __restrict__ phi,
                                                        const double* __restrict__ phib, double* __restrict__ out) {
    int64_t lo, hi;
    block_range(nc, lo, hi);
    for (int64_t c = lo + threadIdx.x; c < hi; c += kBlock) {
        double r[DIM];
#pragma unroll
        for (int a = 0; a < DIM; ++a) r[a] = 0.0;
        for (int32_t j = iptr[c]; j < iptr[c + 1]; ++j) {
            const int64_t f = iidx[j];
            const double coef = wl[f] * (phi[N[f]] - phi[P[f]]);
#pragma unroll
            for (int a = 0; a < DIM; ++a) r[a] += coef * d[f * DIM + a];
        }
        const double pc = phi[c];
        for (int32_t j = bptr[c]; j < bptr[c + 1]; ++j) {
            const int64_t f = bidx[j];
            const double coef = wlb[f] * (phib[f] - pc);
#pragma unroll
            for (int a = 0; a < DIM; ++a) r[a] += coef * db[f * DIM + a];
        }
        const double* G = Ginv + c * DIM * DIM;
#pragma unroll
        for (int a = 0; a < DIM; ++a) {
            double s = 0.0;
#pragma unroll
            for (int b = 0; b < DIM; ++b) s += G[a * DIM + b] * r[b];
            out[c * DIM + a] = s;
        }
    }
}

template <int DIM>
__global__ void __launch_bounds__(kBlock) k_gg_gradient(int64_t nc, const int32_t* __restrict__ iptr,
                                                        const int32_t* __restrict__ iidx,
                                                        const double* __restrict__ isgn,
                                                        const int32_t* __restrict__ bptr,
                                                        const int32_t* __restrict__ bidx,
                                                        const int32_t* __restrict__ P, const int32_t* __restrict__ N,
                                                        const double* __restrict__ w, const double* __restrict__ S,
                                                        const double* __restrict__ Sb, const double* __restrict__ vol,
                                                        const double* __restrict__ dxm,
                                                        const double* __restrict__ phi,
                                                        const double* __restrict__ phib,
                                                        const double* __restrict__ grad, double* __restrict__ out) {
    int64_t lo, hi;
    block_range(nc, lo, hi);
    for (int64_t c = lo + threadIdx.x; c < hi; c += kBlock) {
        double s[DIM];
#pragma unroll
        for (int a = 0; a < DIM; ++a) s[a] = 0.0;
        for (int32_t j = iptr[c]; j < iptr[c + 1]; ++j) {
            const int64_t f = iidx[j];
            const int64_t p = P[f], q = N[f];
            const double wf = w[f], wn = 1.0 - wf;
            double pf = wf * phi[p] + wn * phi[q];
            if (grad) {
#pragma unroll
                for (int a = 0; a < DIM; ++a) pf += (wf * grad[p * DIM + a] + wn * grad[q * DIM + a]) * dxm[f * DIM + a];
            }
            const double coef = isgn[j] * pf;
#pragma unroll
            for (int a = 0; a < DIM; ++a) s[a] += coef * S[f * DIM + a];
        }
        for (int32_t j = bptr[c]; j < bptr[c + 1]; ++j) {
            const int64_t f = bidx[j];
#pragma unroll
            for (int a = 0; a < DIM; ++a) s[a] += phib[f] * Sb[f * DIM + a];
        }
        const double iv = 1.0 / vol[c];
#pragma unroll
        for (int a = 0; a < DIM; ++a) out[c * DIM + a] = s[a] * iv;
    }
}

// CSR data from (diag, aPN, aNP) through the face slots, row by row: the slots of one row
// are written only by that row's thread, and a row's face contributions are added in face
// order, the order of the reference's sequential loop.
__global__ void __launch_bounds__(kBlock) k_assemble(int64_t nc, const int32_t* __restrict__ ptr,
                                                     const int32_t* __restrict__ slot_diag,
                                                     const int32_t* __restrict__ src_ptr,
                                                     const int32_t* __restrict__ src,
                                                     const double* __restrict__ diag, const double* __restrict__ aPN,
                                                     const double* __restrict__ aNP, double* __restrict__ data) {
    int64_t lo, hi;
    block_range(nc, lo, hi);
    for (int64_t c = lo + threadIdx.x; c < hi; c += kBlock) {
        for (int32_t j = ptr[c]; j < ptr[c + 1]; ++j) data[j] = 0.0;
        data[slot_diag[c]] = diag[c];
        // src entries of row c: (face << 1 | which) in face order; which 0 = aPN, 1 = aNP
        for (int32_t e = src_ptr[c]; e < src_ptr[c + 1]; e += 2) {
            const int32_t code = src[e], slot = src[e + 1];
            const int32_t f = code >> 1;
            data[slot] += (code & 1) ? aNP[f] : aPN[f];
        }
    }
}

}  // namespace

extern "C" {

int ftmi_csr_spmm(int64_t nrows, int64_t k, const int32_t* indptr, const int32_t* indices, const double* data,
                  const double* X, double* Y, void* stream) {
    if (nrows <= 0) return 0;
    hipStream_t s = (hipStream_t)stream;
    const int g = grid_for(nrows);
    if (k == 1)
        k_csr_spmm<1><<<g, kBlock, 0, s>>>(nrows, indptr, indices, data, X, Y, 1);
    else
        k_csr_spmm<0><<<g, kBlock, 0, s>>>(nrows, indptr, indices, data, X, Y, (int)k);
    FTMI_LAUNCH_CHECK("ftmi_csr_spmm");
    return 0;
}

int ftmi_cast_f64_f32(int64_t n, const double* x, float* y, void* stream) {
    if (n > 0) k_cast_f32<<<grid_for(n), kBlock, 0, (hipStream_t)stream>>>(n, x, y);
    FTMI_LAUNCH_CHECK("ftmi_cast_f64_f32");
    return 0;
}

int ftmi_ldu_slice(int64_t nhe, const int32_t* pos, const double* off, double* soff, void* stream) {
    if (nhe > 0) k_slice<double><<<grid_for(nhe), kBlock, 0, (hipStream_t)stream>>>(nhe, pos, off, soff);
    FTMI_LAUNCH_CHECK("ftmi_ldu_slice");
    return 0;
}

int ftmi_ldu_slice_f32(int64_t nhe, const int32_t* pos, const float* off, float* soff, void* stream) {
    if (nhe > 0) k_slice<float><<<grid_for(nhe), kBlock, 0, (hipStream_t)stream>>>(nhe, pos, off, soff);
    FTMI_LAUNCH_CHECK("ftmi_ldu_slice_f32");
    return 0;
}

int ftmi_gather_f64(int64_t n, const int64_t* idx, const double* x, double* y, void* stream) {
    if (n > 0) k_gather_f64<<<grid_for(n), kBlock, 0, (hipStream_t)stream>>>(n, idx, x, y);
    FTMI_LAUNCH_CHECK("ftmi_gather_f64");
    return 0;
}

int ftmi_face_interp(int64_t nf, int64_t k, const double* w, const int32_t* P, const int32_t* N, const double* phi,
                     double* out, void* stream) {
    if (nf <= 0) return 0;
    k_face_interp<<<grid_for(nf * k), kBlock, 0, (hipStream_t)stream>>>(nf, (int)k, w, P, N, phi, out);
    FTMI_LAUNCH_CHECK("ftmi_face_interp");
    return 0;
}

int ftmi_ls_gradient(int64_t nc, int64_t dim, const int32_t* iptr, const int32_t* iidx, const int32_t* bptr,
                     const int32_t* bidx, const int32_t* P, const int32_t* N, const double* wl, const double* d,
                     const double* wlb, const double* db, const double* Ginv, const double* phi, const double* phib,
                     double* out, void* stream) {
    if (nc <= 0) return 0;
    hipStream_t s = (hipStream_t)stream;
    if (dim == 2)
        k_ls_gradient<2><<<grid_for(nc), kBlock, 0, s>>>(nc, iptr, iidx, bptr, bidx, P, N, wl, d, wlb, db, Ginv, phi,
                                                          phib, out);
    else if (dim == 3)
        k_ls_gradient<3><<<grid_for(nc), kBlock, 0, s>>>(nc, iptr, iidx, bptr, bidx, P, N, wl, d, wlb, db, Ginv, phi,
                                                          phib, out);
    else {
        set_error_msg("ftmi_ls_gradient: dim must be 2 or 3");
        return -1;
    }
    FTMI_LAUNCH_CHECK("ftmi_ls_gradient");
    return 0;
}

int ftmi_gg_gradient(int64_t nc, int64_t dim, const int32_t* iptr, const int32_t* iidx, const double* isgn,
                     const int32_t* bptr, const int32_t* bidx, const int32_t* P, const int32_t* N, const double* w,
                     const double* S, const double* Sb, const double* vol, const double* dxm, const double* phi,
                     const double* phib, const double* grad, double* out, void* stream) {
    if (nc <= 0) return 0;
    hipStream_t s = (hipStream_t)stream;
    if (dim == 2)
        k_gg_gradient<2><<<grid_for(nc), kBlock, 0, s>>>(nc, iptr, iidx, isgn, bptr, bidx, P, N, w, S, Sb, vol, dxm,
                                                          phi, phib, grad, out);
    else if (dim == 3)
        k_gg_gradient<3><<<grid_for(nc), kBlock, 0, s>>>(nc, iptr, iidx, isgn, bptr, bidx, P, N, w, S, Sb, vol, dxm,
                                                          phi, phib, grad, out);
    else {
        set_error_msg("ftmi_gg_gradient: dim must be 2 or 3");
        return -1;
    }
    FTMI_LAUNCH_CHECK("ftmi_gg_gradient");
    return 0;
}

int ftmi_assemble(int64_t nnz, int64_t nc, int64_t nf, const int32_t* indptr, const int32_t* slot_diag,
                  const int32_t* src_ptr, const int32_t* src, const double* diag, const double* aPN,
                  const double* aNP, double* data, void* stream) {
    (void)nnz;
    (void)nf;
    if (nc <= 0) return 0;
    k_assemble<<<grid_for(nc), kBlock, 0, (hipStream_t)stream>>>(nc, indptr, slot_diag, src_ptr, src, diag, aPN, aNP,
                                                                  data);
    FTMI_LAUNCH_CHECK("ftmi_assemble");
    return 0;
}

}  // extern "C"

// ------------------------------------------------------------------ segmented sums
// out[r * nseg + k] = sum over i in [off[k], off[k+1]) of x[r * stride + i] (|x| when absval):
// one workgroup per (segment, row), strided per thread then a fixed tree, so the sums are the
// same on every run. The per-case residual sums of a batch of cases on one concatenated mesh.
namespace {
__global__ void __launch_bounds__(kBlock) k_segment_sums(int64_t nseg, const int64_t* __restrict__ off,
                                                         int64_t stride, const double* __restrict__ x, int absval,
                                                         double* __restrict__ out) {
    const int64_t k = blockIdx.x, r = blockIdx.y;
    const double* row = x + r * stride;
    double acc[1] = {0.0};
    const int64_t lo = off[k], hi = off[k + 1];
    for (int64_t i = lo + threadIdx.x; i < hi; i += kBlock) acc[0] += absval ? fabs(row[i]) : row[i];
    block_sum<1>(acc);
    if (threadIdx.x == 0) out[r * nseg + k] = acc[0];
}
}  // namespace

extern "C" int ftmi_segment_sums(int64_t nseg, const int64_t* offsets, int64_t nrows, int64_t stride,
                                 const double* x, int absval, double* out, void* stream) {
    if (nseg <= 0 || nrows <= 0) return 0;
    if (nseg > 2147483647 || nrows > 65535) {
        set_error_msg("ftmi_segment_sums: too many segments or rows");
        return -1;
    }
    k_segment_sums<<<dim3((unsigned)nseg, (unsigned)nrows), kBlock, 0, (hipStream_t)stream>>>(nseg, offsets, stride, x,
                                                                                            absval, out);
    FTMI_LAUNCH_CHECK("ftmi_segment_sums");
    return 0;
}
