// Device reductions behind the outer-iteration diagnostics (flowtrace/unstructured/diagnostics.py:1-360).
//
// The diagnostics evaluate the systems an outer iteration assembled and solved, on demand:
//  * per row of an LDU system A x = b at x: r = b - A x and s = |b| + |A| |x| (the
//    Oettli-Prager scale), with b optionally given as b_r - c D x_b (the unrelaxed right-hand
//    side recovered from the relaxed one, b = b_r - (1 - alpha)/alpha a_P u^n);
//  * per cell: the net outflow m = sum of outward face fluxes and the throughflow t = sum |F|;
//  * per row set (several (r, s) vectors sharing one reference scale s_ref = the largest finite
//    s of the set): rows with s <= gamma_K s_ref are excluded, and the metrics are the largest
//    |r|/s over active rows with its (first) row, the active / excluded / non-finite counts, the
//    largest excluded |r|, and the l1, max and l2 of |r| over finite rows;
//  * per vector (or row-vector field): largest row norm of x - y and of x, min, max, non-finite
//    count, sum, sum |x| and sum x^2.
// Every kernel writes its results into a caller buffer of doubles on the device; the caller
// collects them and moves them to the host in one transfer. Reductions are fixed-order
// (workgroup partials, one finalising workgroup), maxima break ties toward the lowest row,
// so the results do not depend on the schedule.
#include <cmath>
#include <cstdint>

#include "../../include/ftmi.h"
#include "ftmi_common.h"

using namespace ftmi;

namespace {

constexpr int kSetMax = 4;   // vectors per row set
constexpr int kStats = 9;    // metrics per vector of a row set
constexpr int kVStats = 8;   // metrics of a vector

struct RowSet {
    const double* r[kSetMax];
    const double* s[kSetMax];
    int ncomp;
};

__global__ void __launch_bounds__(kBlock) k_diag_rows(int64_t n, const double* __restrict__ dg,
                                                      const int32_t* __restrict__ ptr, const int32_t* __restrict__ nbr,
                                                      const double* __restrict__ off, const double* __restrict__ b,
                                                      const double* __restrict__ dgb, double c,
                                                      const double* __restrict__ xb, const double* __restrict__ x,
                                                      double* __restrict__ r, double* __restrict__ s) {
    int64_t lo, hi;
    block_range(n, lo, hi);
    for (int64_t i = lo + threadIdx.x; i < hi; i += kBlock) {
        const double bi = dgb ? b[i] - c * dgb[i] * xb[i] : b[i];
        double ax = dg[i] * x[i];
        double sx = fabs(dg[i]) * fabs(x[i]);
        for (int32_t k = ptr[i]; k < ptr[i + 1]; ++k) {
            const double xn = x[nbr[k]];
            ax += off[k] * xn;
            sx += fabs(off[k]) * fabs(xn);
        }
        r[i] = bi - ax;
        s[i] = fabs(bi) + sx;
    }
}

// m_i = sum over the cell's faces of the outward flux, t_i = sum of |flux|; interior half-edge
// code cface = f << 1 | (cell is the neighbour of f)
__global__ void __launch_bounds__(kBlock) k_diag_flux(int64_t n, const int32_t* __restrict__ cptr,
                                                      const int32_t* __restrict__ cface,
                                                      const int32_t* __restrict__ bptr,
                                                      const int32_t* __restrict__ bface,
                                                      const double* __restrict__ F, const double* __restrict__ Fb,
                                                      double* __restrict__ m, double* __restrict__ t) {
    int64_t lo, hi;
    block_range(n, lo, hi);
    for (int64_t i = lo + threadIdx.x; i < hi; i += kBlock) {
        double mi = 0.0, ti = 0.0;
        for (int32_t k = cptr[i]; k < cptr[i + 1]; ++k) {
            const int32_t code = cface[k];
            const double f = F[code >> 1];
            mi += (code & 1) ? -f : f;
            ti += fabs(f);
        }
        for (int32_t k = bptr[i]; k < bptr[i + 1]; ++k) {
            const double f = Fb[bface[k]];
            mi += f;
            ti += fabs(f);
        }
        m[i] = mi;
        t[i] = ti;
    }
}

__device__ inline bool finite(double v) { return isfinite(v); }

// workgroup partial of the largest finite s over the set
__global__ void __launch_bounds__(kBlock) k_diag_sref(int64_t n, RowSet S, double* __restrict__ part) {
    __shared__ double red[kBlock];
    int64_t lo, hi;
    block_range(n, lo, hi);
    double mx = 0.0;
    for (int c = 0; c < S.ncomp; ++c)
        for (int64_t i = lo + threadIdx.x; i < hi; i += kBlock) {
            const double v = S.s[c][i];
            if (finite(v) && v > mx) mx = v;
        }
    red[threadIdx.x] = mx;
    __syncthreads();
    for (int w = kBlock / 2; w > 0; w >>= 1) {
        if (threadIdx.x < w) red[threadIdx.x] = fmax(red[threadIdx.x], red[threadIdx.x + w]);
        __syncthreads();
    }
    if (threadIdx.x == 0) part[blockIdx.x] = red[0];
}

// the set's statistics, workgroup partials part[block][c][kStats]:
// 0 max q, 1 its row, 2 active, 3 excluded, 4 excluded max |r|, 5 non-finite, 6 l1, 7 max |r|, 8 sum r^2
__global__ void __launch_bounds__(kBlock) k_diag_stats(int64_t n, RowSet S, const double* __restrict__ part_ref,
                                                       int np_ref, double gamma_k, double* __restrict__ part) {
    __shared__ double red[kBlock];
    __shared__ double ridx[kBlock];
    __shared__ double thr_s;
    if (threadIdx.x == 0) {
        double m = 0.0;
        for (int b = 0; b < np_ref; ++b) m = fmax(m, part_ref[b]);
        thr_s = gamma_k * m;
    }
    __syncthreads();
    const double thr = thr_s;
    int64_t lo, hi;
    block_range(n, lo, hi);
    for (int c = 0; c < S.ncomp; ++c) {
        double qmax = 0.0, qrow = -1.0, act = 0.0, exc = 0.0, excmax = 0.0, nonf = 0.0, l1 = 0.0, amax = 0.0, l2 = 0.0;
        for (int64_t i = lo + threadIdx.x; i < hi; i += kBlock) {
            const double r = S.r[c][i], s = S.s[c][i];
            const double ar = fabs(r);
            if (!(finite(r) && finite(s))) {
                nonf += 1.0;
                continue;
            }
            l1 += ar;
            amax = fmax(amax, ar);
            l2 += ar * ar;
            if (s > thr) {
                act += 1.0;
                const double q = ar / s;
                if (qrow < 0.0 || q > qmax) {  // rows ascend per thread: keeps the first maximum
                    qmax = q;
                    qrow = (double)i;
                }
            } else {
                exc += 1.0;
                excmax = fmax(excmax, ar);
            }
        }
        // max q with the lowest row among ties (rows with no active entry carry row -1)
        red[threadIdx.x] = qmax;
        ridx[threadIdx.x] = qrow;
        __syncthreads();
        for (int w = kBlock / 2; w > 0; w >>= 1) {
            if (threadIdx.x < w) {
                const double qa = red[threadIdx.x], qb = red[threadIdx.x + w];
                const double ia = ridx[threadIdx.x], ib = ridx[threadIdx.x + w];
                const bool take_b = ib >= 0.0 && (ia < 0.0 || qb > qa || (qb == qa && ib < ia));
                if (take_b) {
                    red[threadIdx.x] = qb;
                    ridx[threadIdx.x] = ib;
                }
            }
            __syncthreads();
        }
        double* out = part + ((int64_t)blockIdx.x * S.ncomp + c) * kStats;
        if (threadIdx.x == 0) {
            out[0] = red[0];
            out[1] = ridx[0];
        }
        __syncthreads();
        double sums[5] = {act, exc, nonf, l1, l2};
        block_sum<5>(sums);
        double mx[2] = {excmax, amax};
        // maxima through the same LDS tree
        for (int j = 0; j < 2; ++j) {
            red[threadIdx.x] = mx[j];
            __syncthreads();
            for (int w = kBlock / 2; w > 0; w >>= 1) {
                if (threadIdx.x < w) red[threadIdx.x] = fmax(red[threadIdx.x], red[threadIdx.x + w]);
                __syncthreads();
            }
            if (threadIdx.x == 0) mx[j] = red[0];
            __syncthreads();
        }
        if (threadIdx.x == 0) {
            out[2] = sums[0];
            out[3] = sums[1];
            out[4] = mx[0];
            out[5] = sums[2];
            out[6] = sums[3];
            out[7] = mx[1];
            out[8] = sums[4];
        }
    }
}

// out[0] = s_ref, out[1 + c * kStats + j]: the set's statistics (row as a double, -1 if none;
// j = 8 is the l2 norm)
__global__ void __launch_bounds__(kBlock) fin_diag_stats(int ncomp, const double* __restrict__ part, int np,
                                                         const double* __restrict__ part_ref, int np_ref,
                                                         double* __restrict__ out) {
    if (threadIdx.x != 0) return;
    double sref = 0.0;
    for (int b = 0; b < np_ref; ++b) sref = fmax(sref, part_ref[b]);
    out[0] = sref;
    for (int c = 0; c < ncomp; ++c) {
        double q = 0.0, row = -1.0, v[kStats] = {0};
        for (int b = 0; b < np; ++b) {
            const double* p = part + ((int64_t)b * ncomp + c) * kStats;
            if (p[1] >= 0.0 && (row < 0.0 || p[0] > q || (p[0] == q && p[1] < row))) {
                q = p[0];
                row = p[1];
            }
            v[2] += p[2];
            v[3] += p[3];
            v[4] = fmax(v[4], p[4]);
            v[5] += p[5];
            v[6] += p[6];
            v[7] = fmax(v[7], p[7]);
            v[8] += p[8];
        }
        double* o = out + 1 + c * kStats;
        o[0] = q;
        o[1] = row;
        for (int j = 2; j < kStats; ++j) o[j] = v[j];
        o[8] = sqrt(v[8]);
    }
}

// vector statistics, partials part[block][kVStats]: 0 max row norm of x - y, 1 max row norm of
// x, 2 min, 3 max, 4 non-finite entries, 5 sum, 6 sum |x|, 7 sum x^2 (rows: ncomp components
// `stride` apart; min / max / sums over finite entries)
__global__ void __launch_bounds__(kBlock) k_diag_vstats(int64_t n, int ncomp, int64_t stride,
                                                        const double* __restrict__ x, const double* __restrict__ y,
                                                        double* __restrict__ part) {
    __shared__ double red[kBlock];
    int64_t lo, hi;
    block_range(n, lo, hi);
    double dmax = 0.0, xmax = 0.0, mn = INFINITY, mxv = -INFINITY, sums[4] = {0.0, 0.0, 0.0, 0.0};
    for (int64_t i = lo + threadIdx.x; i < hi; i += kBlock) {
        double d2 = 0.0, x2 = 0.0;
        for (int c = 0; c < ncomp; ++c) {
            const double v = x[c * stride + i];
            if (!finite(v)) {
                sums[0] += 1.0;
                continue;
            }
            x2 += v * v;
            mn = fmin(mn, v);
            mxv = fmax(mxv, v);
            sums[1] += v;
            sums[2] += fabs(v);
            sums[3] += v * v;
            if (y) {
                const double dv = v - y[c * stride + i];
                d2 += dv * dv;
            }
        }
        dmax = fmax(dmax, sqrt(d2));
        xmax = fmax(xmax, sqrt(x2));
    }
    block_sum<4>(sums);
    double mx[4] = {dmax, xmax, -mn, mxv};
    for (int j = 0; j < 4; ++j) {
        red[threadIdx.x] = mx[j];
        __syncthreads();
        for (int w = kBlock / 2; w > 0; w >>= 1) {
            if (threadIdx.x < w) red[threadIdx.x] = fmax(red[threadIdx.x], red[threadIdx.x + w]);
            __syncthreads();
        }
        if (threadIdx.x == 0) mx[j] = red[0];
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        double* o = part + (int64_t)blockIdx.x * kVStats;
        o[0] = mx[0];
        o[1] = mx[1];
        o[2] = -mx[2];
        o[3] = mx[3];
        o[4] = sums[0];
        o[5] = sums[1];
        o[6] = sums[2];
        o[7] = sums[3];
    }
}

__global__ void __launch_bounds__(kBlock) fin_diag_vstats(const double* __restrict__ part, int np,
                                                          double* __restrict__ out) {
    if (threadIdx.x != 0) return;
    double v[kVStats] = {0.0, 0.0, INFINITY, -INFINITY, 0.0, 0.0, 0.0, 0.0};
    for (int b = 0; b < np; ++b) {
        const double* p = part + (int64_t)b * kVStats;
        v[0] = fmax(v[0], p[0]);
        v[1] = fmax(v[1], p[1]);
        v[2] = fmin(v[2], p[2]);
        v[3] = fmax(v[3], p[3]);
        for (int j = 4; j < kVStats; ++j) v[j] += p[j];
    }
    for (int j = 0; j < kVStats; ++j) out[j] = v[j];
}

int diag_grid(int64_t n) { return grid_red(n > 0 ? n : 1); }

}  // namespace

extern "C" {

int ftmi_diag_rows(int64_t n, const double* dg, const int32_t* ptr, const int32_t* nbr, const double* off,
                   const double* b, const double* dgb, double c, const double* xb, const double* x, double* r,
                   double* s, void* stream) {
    if (n <= 0) return 0;
    if (dgb && !xb) {
        set_error_msg("ftmi_diag_rows: dgb needs xb");
        return -1;
    }
    k_diag_rows<<<grid_for(n), kBlock, 0, (hipStream_t)stream>>>(n, dg, ptr, nbr, off, b, dgb, c, xb, x, r, s);
    FTMI_LAUNCH_CHECK("ftmi_diag_rows");
    return 0;
}

int ftmi_diag_cell_flux(int64_t n, const int32_t* cptr, const int32_t* cface, const int32_t* bptr,
                        const int32_t* bface, const double* F, const double* Fb, double* m, double* t, void* stream) {
    if (n <= 0) return 0;
    k_diag_flux<<<grid_for(n), kBlock, 0, (hipStream_t)stream>>>(n, cptr, cface, bptr, bface, F, Fb, m, t);
    FTMI_LAUNCH_CHECK("ftmi_diag_cell_flux");
    return 0;
}

// doubles of work that ftmi_diag_rowset and ftmi_diag_vstats need
int64_t ftmi_diag_work_doubles(void) { return (int64_t)kRedGrid * (1 + kSetMax * kStats); }

int ftmi_diag_rowset(int64_t n, int ncomp, const double* const* r, const double* const* s, double gamma_k,
                     double* out, double* work, void* stream) {
    if (ncomp < 1 || ncomp > kSetMax) {
        set_error_msg("ftmi_diag_rowset: 1 to 4 vectors per set");
        return -1;
    }
    RowSet S{};
    S.ncomp = ncomp;
    for (int c = 0; c < ncomp; ++c) {
        S.r[c] = r[c];
        S.s[c] = s[c];
    }
    hipStream_t st = (hipStream_t)stream;
    const int g = diag_grid(n);
    double* part_ref = work;
    double* part = work + kRedGrid;
    k_diag_sref<<<g, kBlock, 0, st>>>(n, S, part_ref);
    k_diag_stats<<<g, kBlock, 0, st>>>(n, S, part_ref, g, gamma_k, part);
    fin_diag_stats<<<1, 64, 0, st>>>(ncomp, part, g, part_ref, g, out);
    FTMI_LAUNCH_CHECK("ftmi_diag_rowset");
    return 0;
}

int ftmi_diag_vstats(int64_t n, int ncomp, int64_t stride, const double* x, const double* y, double* out,
                     double* work, void* stream) {
    if (ncomp < 1) {
        set_error_msg("ftmi_diag_vstats: ncomp >= 1");
        return -1;
    }
    hipStream_t st = (hipStream_t)stream;
    const int g = diag_grid(n);
    k_diag_vstats<<<g, kBlock, 0, st>>>(n, ncomp, stride, x, y, work);
    fin_diag_vstats<<<1, 64, 0, st>>>(work, g, out);
    FTMI_LAUNCH_CHECK("ftmi_diag_vstats");
    return 0;
}

}  // extern "C"
