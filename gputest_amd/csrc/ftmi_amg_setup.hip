// Smoothed-aggregation AMG set-up on the GPU (flowtrace/linalg/amg.py:60-117).
//
// Everything but the greedy aggregation runs here: CSR of the pressure matrix from its face
// form, the diagonal, the strength mask |a_ij| >= theta sqrt(|a_ii a_jj|), the tentative
// prolongator weights 1/sqrt(|aggregate|), rho(D^-1 A) by power iteration, the smoothed
// prolongator P = T - 4/(3 rho) D^-1 A T, R = P^T, and the Galerkin product R (A P). The
// aggregation itself (three greedy passes whose result depends on the node order) runs on
// the host over the strength mask copied back (ftmi_amg_aggregate_mask_host).
//
// Sparse products run one thread per output row as a k-way merge of the right operand's rows
// (rows of the left operand up to 32 entries); longer rows go to one wavefront each, whose
// lanes insert the row's candidate columns into an LDS hash table (integer compare-and-swap,
// so the set is deterministic), accumulate staged products in LDS and sort the distinct
// columns with a bitonic network. Either way
// each column's value is summed in the fixed order of the left operand's row, so results are
// independent of the schedule and equal a row-by-row sparse product. Rows are sized in a
// count pass, the caller scans the counts into row pointers and allocates, and a fill pass
// writes the entries.
#include <cstdio>
#include <cstring>

#include "../../include/ftmi.h"
#include "ftmi_common.h"
#include "ftmi_ops.h"

using namespace ftmi;

namespace {

constexpr int kHash = 2048;     // LDS hash slots per row (distinct columns must stay below)
constexpr int kHashLimit = 1536;

__global__ void __launch_bounds__(kBlock) k_ldu_to_csr(int64_t n, const int32_t* __restrict__ cptr,
                                                       const int32_t* __restrict__ he_slot,
                                                       const int32_t* __restrict__ slot_diag,
                                                       const int32_t* __restrict__ indptr,
                                                       const double* __restrict__ diag,
                                                       const double* __restrict__ off, double* __restrict__ data) {
    int64_t lo, hi;
    block_range(n, lo, hi);
    for (int64_t c = lo + threadIdx.x; c < hi; c += kBlock) {
        for (int32_t j = indptr[c]; j < indptr[c + 1]; ++j) data[j] = 0.0;
        data[slot_diag[c]] = diag[c];
        for (int32_t k = cptr[c]; k < cptr[c + 1]; ++k) data[he_slot[k]] += off[k];
    }
}

// Row r is handled by TPR lanes (1 on the sparse fine levels, 64 on dense coarse levels).
template <int TPR>
struct LaneRows {
    int64_t first, stride;
    int lane;
    __device__ LaneRows() {
        const int g = gridDim.x, b = blockIdx.x;
        int logical = b;
        if (g >= kNXcd && g % kNXcd == 0) logical = (b % kNXcd) * (g / kNXcd) + b / kNXcd;
        first = (int64_t)logical * (kBlock / TPR) + threadIdx.x / TPR;
        stride = (int64_t)g * (kBlock / TPR);
        lane = threadIdx.x % TPR;
    }
};

// Lanes per row: the V-cycle's choice (entries per row <= 8: 1 lane, the plain column-order
// sum; <= 24: 4; <= 96: 16; else 64), so that the long rows of the coarse Galerkin levels are
// read by several lanes at once instead of one lane walking hundreds of entries.
inline int rows_tpr(int64_t n, int64_t nnz) {
    const double avg = n > 0 ? (double)nnz / (double)n : 0.0;
    if (avg <= 8.0) return 1;
    if (avg <= 24.0) return 4;
    if (avg <= 96.0) return 16;
    return 64;
}
inline int rows_grid(int64_t n, int tpr) { return grid_for(n * tpr); }

#define ROWS_TPR_LAUNCH(tpr, KERNEL, grid, st, ...)                        \
    do {                                                                   \
        switch (tpr) {                                                     \
        case 64: KERNEL<64><<<(grid), kBlock, 0, (st)>>>(__VA_ARGS__); break; \
        case 16: KERNEL<16><<<(grid), kBlock, 0, (st)>>>(__VA_ARGS__); break; \
        case 4: KERNEL<4><<<(grid), kBlock, 0, (st)>>>(__VA_ARGS__); break;   \
        default: KERNEL<1><<<(grid), kBlock, 0, (st)>>>(__VA_ARGS__); break;  \
        }                                                                  \
    } while (0)

// the diagonal entry (the row's one entry with col == row, else 0)
template <int TPR>
__global__ void __launch_bounds__(kBlock) k_diag(int64_t n, const int32_t* __restrict__ ptr,
                                                 const int32_t* __restrict__ col, const double* __restrict__ val,
                                                 double* __restrict__ diag, double* __restrict__ dinv) {
    LaneRows<TPR> R;
    for (int64_t i = R.first; i < n; i += R.stride) {
        double d = 0.0;
        for (int32_t j = ptr[i] + R.lane; j < ptr[i + 1]; j += TPR)
            if (col[j] == i) d += val[j];
#pragma unroll
        for (int off = TPR / 2; off > 0; off >>= 1) d += __shfl_xor(d, off, TPR);
        if (R.lane == 0) {
            diag[i] = d;
            dinv[i] = 1.0 / d;
        }
    }
}

// With old != nullptr the new mask is also compared with old entry by entry, and a wave that
// finds a difference stores gen into *flag (every such store writes the same value; the
// caller passes a gen that *flag does not hold yet, so no clearing pass is needed).
template <int TPR>
__global__ void __launch_bounds__(kBlock) k_strength(int64_t n, const int32_t* __restrict__ ptr,
                                                     const int32_t* __restrict__ col,
                                                     const double* __restrict__ val,
                                                     const double* __restrict__ diag, double theta,
                                                     uint8_t* __restrict__ mask,
                                                     const uint8_t* __restrict__ old, int32_t* flag,
                                                     int32_t gen) {
    LaneRows<TPR> R;
    bool differs = false;
    for (int64_t i = R.first; i < n; i += R.stride) {
        const double di = fabs(diag[i]);
        for (int32_t j = ptr[i] + R.lane; j < ptr[i + 1]; j += TPR) {
            const int32_t c = col[j];
            const uint8_t m = (c != i && fabs(val[j]) >= theta * sqrt(di * fabs(diag[c]))) ? 1 : 0;
            mask[j] = m;
            if (old != nullptr && old[j] != m) differs = true;
        }
    }
    if (old != nullptr && __ballot(differs) != 0ull && (threadIdx.x & 63) == 0) *flag = gen;
}

// *flag = gen when the n 64-bit words of a and b differ anywhere (bitwise, so -0.0 and NaN
// payloads count as differences): "is this exactly the matrix a hierarchy was built for?"
__global__ void __launch_bounds__(kBlock) k_differs_u64(int64_t n, const uint64_t* __restrict__ a,
                                                        const uint64_t* __restrict__ b, int32_t* flag,
                                                        int32_t gen) {
    bool differs = false;
    for (int64_t i = blockIdx.x * (int64_t)kBlock + threadIdx.x; i < n; i += (int64_t)gridDim.x * kBlock)
        if (a[i] != b[i]) differs = true;
    if (__ballot(differs) != 0ull && (threadIdx.x & 63) == 0) *flag = gen;
}

// the strength graph as its own CSR (the strong entries of each row in column order): row
// counts, then (after a scan of the counts) the columns
__global__ void __launch_bounds__(kBlock) k_strong_count(int64_t n, const int32_t* __restrict__ ptr,
                                                         const uint8_t* __restrict__ mask, int32_t* __restrict__ cnt) {
    int64_t lo, hi;
    block_range(n, lo, hi);
    for (int64_t i = lo + threadIdx.x; i < hi; i += kBlock) {
        int32_t c = 0;
        for (int32_t j = ptr[i]; j < ptr[i + 1]; ++j) c += mask[j];
        cnt[i] = c;
    }
}

__global__ void __launch_bounds__(kBlock) k_strong_fill(int64_t n, const int32_t* __restrict__ ptr,
                                                        const int32_t* __restrict__ col,
                                                        const uint8_t* __restrict__ mask,
                                                        const int32_t* __restrict__ s_ptr,
                                                        int32_t* __restrict__ s_col) {
    int64_t lo, hi;
    block_range(n, lo, hi);
    for (int64_t i = lo + threadIdx.x; i < hi; i += kBlock) {
        int32_t q = s_ptr[i];
        for (int32_t j = ptr[i]; j < ptr[i + 1]; ++j)
            if (mask[j]) s_col[q++] = col[j];
    }
}

// T as a CSR matrix with one entry per row: (agg[i], 1/sqrt(size of aggregate agg[i]))
__global__ void __launch_bounds__(kBlock) k_tentative(int64_t n, const int32_t* __restrict__ agg,
                                                      const int32_t* __restrict__ agg_size,
                                                      int32_t* __restrict__ t_ptr, double* __restrict__ t_val) {
    int64_t lo, hi;
    block_range(n + 1, lo, hi);
    for (int64_t i = lo + threadIdx.x; i < hi; i += kBlock) {
        t_ptr[i] = (int32_t)i;
        if (i < n) t_val[i] = 1.0 / sqrt((double)agg_size[agg[i]]);
    }
}

// ---------------------------------------------------------------- power iteration for rho(D^-1 A)
// One launch per iteration: the previous iterate is normalised on the fly (x = y_prev /
// max(||y_prev||, 1e-300), ||y_prev|| re-added from the previous partials by every
// workgroup), then y = (D^-1 A) x with the products (d_i a_ij) x_j of a row-by-row sparse
// product, and the partials of ||y||^2 and ||x||^2. The first iteration takes x as given.
// Rows are handled by TPR lanes (rows_tpr): 1 keeps the column-order sum, wider groups sum in a
// fixed lane tree. The first iteration is a kernel of its own (k_pow_first): a run-time flag
// put a branch, and a wait for its load, inside every gather of a row.
template <int TPR, bool FIRST>
__device__ inline void pow_step(int64_t n, const int32_t* __restrict__ ptr, const int32_t* __restrict__ col,
                                const double* __restrict__ val, const double* __restrict__ dinv,
                                const double* __restrict__ y_prev, double* __restrict__ x, double* __restrict__ y,
                                const double* part_in, int np_in, double* part_out) {
    double scale = 1.0;
    if (!FIRST) {
        double tot[2];
        consume_sum<2>(part_in, np_in, tot);
        const double ny = sqrt(tot[0]);
        scale = ny > 1e-300 ? ny : 1e-300;
    }
    auto xv = [&](int64_t j) {
        if (FIRST) return x[j];
        return y_prev[j] / scale;
    };
    const int g = gridDim.x, b = blockIdx.x;
    int logical = b;
    if (g >= kNXcd && g % kNXcd == 0) logical = (b % kNXcd) * (g / kNXcd) + b / kNXcd;
    const int64_t stride = (int64_t)g * (kBlock / TPR);
    const int lane = threadIdx.x % TPR;
    double acc[2] = {0.0, 0.0};
    for (int64_t i = (int64_t)logical * (kBlock / TPR) + threadIdx.x / TPR; i < n; i += stride) {
        const double di = dinv[i];
        // this lane's entries lane, lane + TPR, ... in order, four loads in flight at a time
        const int32_t p0 = ptr[i] + lane, e = ptr[i + 1];
        const int d = p0 < e ? (e - p0 + TPR - 1) / TPR : 0;
        double s = gather_sum(
            0.0, d, (int32_t)i, [&](int k) { return col[p0 + TPR * k]; },
            [&](int k) { return di * val[p0 + TPR * k]; }, [&](int32_t j) { return xv(j); });
#pragma unroll
        for (int off = TPR / 2; off > 0; off >>= 1) s += __shfl_xor(s, off, TPR);
        if (lane == 0) {
            const double xi = xv(i);
            if (!FIRST) x[i] = xi;
            y[i] = s;
            acc[0] += s * s;
            acc[1] += xi * xi;
        }
    }
    block_partials<2>(acc, part_out);
}

template <int TPR>
__global__ void __launch_bounds__(kBlock) k_pow_first(int64_t n, const int32_t* __restrict__ ptr,
                                                      const int32_t* __restrict__ col, const double* __restrict__ val,
                                                      const double* __restrict__ dinv, const double* __restrict__ y_prev,
                                                      double* __restrict__ x, double* __restrict__ y,
                                                      const double* part_in, int np_in, double* part_out) {
    pow_step<TPR, true>(n, ptr, col, val, dinv, y_prev, x, y, part_in, np_in, part_out);
}

template <int TPR>
__global__ void __launch_bounds__(kBlock) k_pow_step(int64_t n, const int32_t* __restrict__ ptr,
                                                     const int32_t* __restrict__ col, const double* __restrict__ val,
                                                     const double* __restrict__ dinv, const double* __restrict__ y_prev,
                                                     double* __restrict__ x, double* __restrict__ y,
                                                     const double* part_in, int np_in, double* part_out) {
    pow_step<TPR, false>(n, ptr, col, val, dinv, y_prev, x, y, part_in, np_in, part_out);
}

// lam = ||y|| / max(||x||, 1e-300)
__global__ void __launch_bounds__(kBlock) k_pow_fin(const double* partials, int nparts, double* lam) {
    double tot[2];
    if (finalize_sum<2>(partials, nparts, tot)) {
        const double ny = sqrt(tot[0]), nx = sqrt(tot[1]);
        *lam = ny / (nx > 1e-300 ? nx : 1e-300);
    }
}

// ---------------------------------------------------------------- wave-per-row sparse products
struct CsrV {
    const int32_t* ptr;
    const int32_t* col;
    const double* val;
};

// multiplicative hash into a table of 2^bits slots
__device__ inline uint32_t hslot(int32_t c, int bits) { return ((uint32_t)c * 2654435761u) >> (32 - bits); }

// position of column c in row k of Y (columns sorted), or -1
__device__ inline int32_t find_in_row(const CsrV& Y, int32_t k, int32_t c) {
    int32_t lo = Y.ptr[k], hi = Y.ptr[k + 1] - 1;
    while (lo <= hi) {
        const int32_t mid = (lo + hi) >> 1;
        const int32_t m = Y.col[mid];
        if (m == c) return mid;
        if (m < c)
            lo = mid + 1;
        else
            hi = mid - 1;
    }
    return -1;
}

// C = Xs Y for the rows listed in rows[0 .. *n_rows) (the rows too long for the merge kernel
// below), one wavefront per row, Xs = diag(xscale) X when xscale != nullptr. In "prolongator"
// mode (agg != nullptr) the row is finished as C_rc = [c == agg[r]] t[r] - omega C_rc, the
// smoothed prolongator T - omega D^-1 A T with Y = T.
//
// The lanes insert the row's candidate columns into an LDS hash table (integer CAS, so the set
// of distinct columns does not depend on the schedule). In the fill pass the candidates'
// products (xs x_rj) y_jc are staged in LDS with the hash slot of their column, a run of
// consecutive j at a time (as many as the staging arrays hold), and added into per-slot
// accumulators one j at a time in the order of X's row (the columns of one Y row are distinct,
// so the lanes of one step never collide): the order a row-by-row sparse product sums in. The
// distinct columns are sorted by a bitonic network and written out with their accumulated
// values. Rows longer than kMaxX, or with a Y row longer than kCand, fall back to summing each
// column over X's row with a binary search in the Y rows (same order, slower).
constexpr int kCand = 2048;  // staged candidates per run
constexpr int kMaxX = 1024;  // longest X row that is staged

// slot of column c in the hash table (c is present)
__device__ inline int find_slot(const int32_t* keys, int32_t c, int bits) {
    uint32_t h = hslot(c, bits);
    const int mask = (1 << bits) - 1;
    while (keys[h] != c) h = (h + 1) & mask;
    return (int)h;
}

template <bool FILL>
__global__ void __launch_bounds__(kWave) k_spgemm(const int32_t* __restrict__ rows, const int32_t* n_rows, CsrV X,
                                                  CsrV Y, const double* __restrict__ xscale,
                                                  const int32_t* __restrict__ agg, const double* __restrict__ tval,
                                                  const double* omega_num, double omega_scale,
                                                  int32_t* __restrict__ rowcnt, const int32_t* __restrict__ cptr,
                                                  int32_t* __restrict__ ccol, double* __restrict__ cval,
                                                  int* overflow) {
    __shared__ int32_t keys[kHash];
    __shared__ int32_t dense[kHash];
    __shared__ double acc[FILL ? kHash : 1];
    __shared__ double stv[FILL ? kCand : 1];
    __shared__ int16_t sts[FILL ? kCand : 1];
    __shared__ int32_t jptr[FILL ? kMaxX + 1 : 1];
    __shared__ int32_t nk, nd;
    const int lane = threadIdx.x;
    const int64_t nlist = *n_rows;
    for (int64_t li = blockIdx.x; li < nlist; li += gridDim.x) {
        const int64_t r = rows[li];
        const int32_t xb = X.ptr[r], xe = X.ptr[r + 1];
        const int32_t L = xe - xb;
        // candidate count, and (fill pass) the offsets of each j's candidates
        int bound = 0, longest = 0;
        const bool stage_ok = FILL && L <= kMaxX;
        if (stage_ok && lane == 0) jptr[0] = 0;
        for (int32_t j0 = 0; j0 < L; j0 += kWave) {
            const int32_t j = j0 + lane;
            int len = 0;
            if (j < L) {
                const int32_t k = X.col[xb + j];
                len = Y.ptr[k + 1] - Y.ptr[k];
            }
            longest = len > longest ? len : longest;
            int inc = len;  // inclusive scan over the wave
#pragma unroll
            for (int off = 1; off < kWave; off <<= 1) {
                const int t = __shfl_up(inc, off, kWave);
                if (lane >= off) inc += t;
            }
            if (stage_ok && j < L) jptr[j + 1] = bound + inc;
            bound += __shfl(inc, kWave - 1, kWave);
        }
#pragma unroll
        for (int off = kWave / 2; off > 0; off >>= 1) {
            const int t = __shfl_xor(longest, off, kWave);
            longest = t > longest ? t : longest;
        }
        const bool staged = stage_ok && longest <= kCand;
        // table size: a power of two >= twice this row's candidate count (at least 64 slots)
        int bits = 6;
        while (bits < 11 && (1 << bits) < 2 * bound) ++bits;
        const int hsize = 1 << bits;
        for (int h = lane; h < hsize; h += kWave) {
            keys[h] = -1;
            if (FILL) acc[h] = 0.0;
        }
        if (lane == 0) {
            nk = 0;
            nd = 0;
        }
        __syncthreads();
        const double xs = (FILL && xscale) ? xscale[r] : 1.0;
        for (int32_t j = xb + lane; j < xe; j += kWave) {
            const int32_t k = X.col[j];
            for (int32_t q = Y.ptr[k]; q < Y.ptr[k + 1]; ++q) {
                const int32_t c = Y.col[q];
                uint32_t h = hslot(c, bits);
                for (int probe = 0; probe < hsize; ++probe) {
                    const int32_t old = atomicCAS(&keys[h], -1, c);
                    if (old == -1) {
                        atomicAdd(&nk, 1);
                        break;
                    }
                    if (old == c) break;
                    h = (h + 1) & (hsize - 1);
                }
            }
        }
        __syncthreads();
        const int n_keys = nk;
        if (n_keys > kHashLimit || n_keys >= hsize) {
            // overflow[0]: rows over capacity, [1]: largest distinct count, [2]: largest
            // candidate bound, [3]: one offending row
            if (lane == 0) {
                atomicAdd(&overflow[0], 1);
                atomicMax(&overflow[1], n_keys);
                atomicMax(&overflow[2], bound);
                overflow[3] = (int)r;
            }
            __syncthreads();
            continue;
        }
        if (!FILL) {
            if (lane == 0) rowcnt[r] = n_keys;
            __syncthreads();
            continue;
        }
        // the distinct columns, compacted (any order) and sorted ascending
        for (int h = lane; h < hsize; h += kWave) {
            const int32_t c = keys[h];
            if (c >= 0) dense[atomicAdd(&nd, 1)] = c;
        }
        int m = kWave;
        while (m < n_keys) m <<= 1;  // <= kHash, as n_keys < hsize <= kHash
        __syncthreads();
        for (int t = n_keys + lane; t < m; t += kWave) dense[t] = 0x7fffffff;
        __syncthreads();
        for (int k = 2; k <= m; k <<= 1) {
            for (int jj = k >> 1; jj > 0; jj >>= 1) {
                for (int i = lane; i < m; i += kWave) {
                    const int ixj = i ^ jj;
                    if (ixj > i) {
                        const int32_t a = dense[i], b = dense[ixj];
                        if ((a > b) == ((i & k) == 0)) {
                            dense[i] = b;
                            dense[ixj] = a;
                        }
                    }
                }
                __syncthreads();
            }
        }
        if (staged) {
            for (int32_t jb = 0; jb < L;) {
                // the run [jb, je): as many j as fit the staging arrays (at least one)
                int32_t je = jb + 1;
                while (je < L && jptr[je + 1] - jptr[jb] <= kCand) ++je;
                const int32_t p0 = jptr[jb];
                for (int32_t j = jb + lane; j < je; j += kWave) {
                    const int32_t k = X.col[xb + j];
                    const int32_t qb = Y.ptr[k], qe = Y.ptr[k + 1];
                    const double xv = xscale ? xs * X.val[xb + j] : X.val[xb + j];
                    const int32_t base = jptr[j] - p0 - qb;
                    for (int32_t q = qb; q < qe; ++q) {
                        sts[base + q] = (int16_t)find_slot(keys, Y.col[q], bits);
                        stv[base + q] = xv * Y.val[q];
                    }
                }
                __syncthreads();
                // one j at a time, in X's order
                for (int32_t j = jb; j < je; ++j) {
                    const int32_t pb = jptr[j] - p0, pe = jptr[j + 1] - p0;
                    for (int32_t p = pb + lane; p < pe; p += kWave) acc[sts[p]] += stv[p];
                    __syncthreads();
                }
                jb = je;
            }
        }
        const int32_t base = cptr[r];
        const double om = agg ? omega_scale / *omega_num : 0.0;  // (4/3) / rho
        for (int t = lane; t < n_keys; t += kWave) {
            const int32_t c = dense[t];
            double v = 0.0;
            if (staged) {
                v = acc[find_slot(keys, c, bits)];
            } else {
                for (int32_t j = xb; j < xe; ++j) {
                    const int32_t q = find_in_row(Y, X.col[j], c);
                    if (q >= 0) v += (xscale ? xs * X.val[j] : X.val[j]) * Y.val[q];
                }
            }
            if (agg) v = (c == agg[r] ? tval[r] : 0.0) - om * v;
            ccol[base + t] = c;
            cval[base + t] = v;
        }
        __syncthreads();
    }
}

// C = Xs Y one wavefront per row when C has few columns (the coarse levels): a dense LDS
// accumulator indexed by column, with one flag bit per column, replaces the hash table and the
// sort. The products are added one j at a time in the order of X's row (the columns of one Y
// row are distinct, so the lanes of one step never collide). The touched columns are written
// out in column order (a wave prefix scan of the flag words' popcounts, then each lane's bits
// in ascending order), and the accumulators and flags they used are cleared on the way, so a
// row costs its candidates plus ncols / 2048 wave steps.
constexpr int kDenseCols = 4096;
constexpr int64_t kDenseWork = int64_t(1) << 24;  // rows x columns up to which the dense kernel is used

template <bool FILL>
__global__ void __launch_bounds__(kWave) k_spgemm_dense(int64_t nrows, int32_t ncols, CsrV X, CsrV Y,
                                                        const double* __restrict__ xscale,
                                                        const int32_t* __restrict__ agg,
                                                        const double* __restrict__ tval, const double* omega_num,
                                                        double omega_scale, int32_t* __restrict__ rowcnt,
                                                        const int32_t* __restrict__ cptr,
                                                        int32_t* __restrict__ ccol, double* __restrict__ cval) {
    extern __shared__ double dense_lds[];  // FILL: ncols accumulators; then the flag words
    double* acc = dense_lds;
    uint32_t* flag = (uint32_t*)(dense_lds + (FILL ? ncols : 0));
    const int32_t nw = (ncols + 31) >> 5;
    const int lane = threadIdx.x;
    const double om = (FILL && agg) ? omega_scale / *omega_num : 0.0;  // (4/3) / rho
    if (FILL)
        for (int32_t c = lane; c < ncols; c += kWave) acc[c] = 0.0;
    for (int32_t w = lane; w < nw; w += kWave) flag[w] = 0u;
    __syncthreads();
    for (int64_t r = blockIdx.x; r < nrows; r += gridDim.x) {
        const int32_t xb = X.ptr[r], xe = X.ptr[r + 1];
        const double xs = (FILL && xscale) ? xscale[r] : 1.0;
        for (int32_t j = xb; j < xe; ++j) {
            const int32_t k = X.col[j];
            const double xv = FILL ? (xscale ? xs * X.val[j] : X.val[j]) : 0.0;
            for (int32_t q = Y.ptr[k] + lane; q < Y.ptr[k + 1]; q += kWave) {
                const int32_t c = Y.col[q];
                atomicOr(&flag[c >> 5], 1u << (c & 31));
                if (FILL) acc[c] += xv * Y.val[q];
            }
            if (FILL) __syncthreads();
        }
        __syncthreads();
        int32_t cnt = 0;
        const int32_t base = FILL ? cptr[r] : 0;
        const int32_t ar = (FILL && agg) ? agg[r] : -1;
        const double tr = (FILL && agg) ? tval[r] : 0.0;
        for (int32_t w0 = 0; w0 < nw; w0 += kWave) {
            const int32_t w = w0 + lane;
            uint32_t word = w < nw ? flag[w] : 0u;
            const int pc = __popc(word);
            int inc = pc;
#pragma unroll
            for (int off = 1; off < kWave; off <<= 1) {
                const int t = __shfl_up(inc, off, kWave);
                if (lane >= off) inc += t;
            }
            if (word) {
                flag[w] = 0u;
                int32_t pos = cnt + inc - pc;
                while (word) {
                    const int32_t c = (w << 5) + (__ffs(word) - 1);
                    word &= word - 1u;
                    if (FILL) {
                        double v = acc[c];
                        acc[c] = 0.0;
                        if (agg) v = (c == ar ? tr : 0.0) - om * v;
                        ccol[base + pos] = c;
                        cval[base + pos] = v;
                    }
                    ++pos;
                }
            }
            cnt += __shfl(inc, kWave - 1, kWave);
        }
        if (!FILL && lane == 0) rowcnt[r] = cnt;
        __syncthreads();
    }
}

// C = Xs Y one thread per row, for rows of X with at most kMergeCap entries: a k-way merge of
// the (column-sorted) rows of Y selected by the row of X. Each step takes the smallest head
// column, adds the contributions of the heads at that column in the order of X's row (the
// order a row-by-row sparse product accumulates in) and advances them, so the row comes out
// sorted with no hash and no LDS. Longer rows, and rows with more than kMergeWork candidates
// (a thread would step through hundreds of output columns alone), are appended to long_rows in
// the count pass and handled by k_spgemm.
constexpr int kMergeCap = 32;
constexpr int kMergeWork = 1024;

template <bool FILL>
__global__ void __launch_bounds__(kBlock) k_spgemm_merge(int64_t nrows, CsrV X, CsrV Y,
                                                         const double* __restrict__ xscale,
                                                         const int32_t* __restrict__ agg,
                                                         const double* __restrict__ tval, const double* omega_num,
                                                         double omega_scale, int32_t* __restrict__ rowcnt,
                                                         const int32_t* __restrict__ cptr,
                                                         int32_t* __restrict__ ccol, double* __restrict__ cval,
                                                         int32_t* __restrict__ long_rows, int32_t* n_long) {
    constexpr int32_t kNone = 0x7fffffff;
    int64_t lo, hi;
    block_range(nrows, lo, hi);
    for (int64_t r = lo + threadIdx.x; r < hi; r += kBlock) {
        const int32_t xb = X.ptr[r];
        const int32_t L = X.ptr[r + 1] - xb;
        if (L > kMergeCap) {
            if (!FILL) long_rows[atomicAdd(n_long, 1)] = (int32_t)r;
            continue;
        }
        int32_t q[kMergeCap], e[kMergeCap], c[kMergeCap];
        int32_t cur = kNone;
        int32_t work = 0;
#pragma unroll
        for (int t = 0; t < kMergeCap; ++t) {
            q[t] = 0;
            e[t] = 0;
            c[t] = kNone;
            if (t < L) {
                const int32_t k = X.col[xb + t];
                q[t] = Y.ptr[k];
                e[t] = Y.ptr[k + 1];
                work += e[t] - q[t];
                if (q[t] < e[t]) c[t] = Y.col[q[t]];
            }
            cur = c[t] < cur ? c[t] : cur;
        }
        if (work > kMergeWork) {
            if (!FILL) long_rows[atomicAdd(n_long, 1)] = (int32_t)r;
            continue;
        }
        const double xs = (FILL && xscale) ? xscale[r] : 1.0;
        const int32_t ar = (FILL && agg) ? agg[r] : -1;
        const double tr = (FILL && agg) ? tval[r] : 0.0;
        const double om = (FILL && agg) ? omega_scale / *omega_num : 0.0;
        const int32_t base = FILL ? cptr[r] : 0;
        int32_t cnt = 0;
        while (cur != kNone) {
            double v = 0.0;
            int32_t nxt = kNone;
#pragma unroll
            for (int t = 0; t < kMergeCap; ++t) {
                if (c[t] == cur) {
                    if (FILL) {
                        const double xv = X.val[xb + t];
                        v += (xscale ? xs * xv : xv) * Y.val[q[t]];
                    }
                    ++q[t];
                    c[t] = q[t] < e[t] ? Y.col[q[t]] : kNone;
                }
                nxt = c[t] < nxt ? c[t] : nxt;
            }
            if (FILL) {
                if (agg) v = (cur == ar ? tr : 0.0) - om * v;
                ccol[base + cnt] = cur;
                cval[base + cnt] = v;
            }
            ++cnt;
            cur = nxt;
        }
        if (!FILL) rowcnt[r] = cnt;
    }
}

// ---------------------------------------------------------------- transpose (R = P^T)
__global__ void __launch_bounds__(kBlock) k_count_cols(int64_t nnz, const int32_t* __restrict__ col,
                                                       int32_t* __restrict__ cnt) {
    int64_t lo, hi;
    block_range(nnz, lo, hi);
    for (int64_t j = lo + threadIdx.x; j < hi; j += kBlock) atomicAdd(&cnt[col[j]], 1);
}

__global__ void __launch_bounds__(kBlock) k_place(int64_t nrows, const int32_t* __restrict__ ptr,
                                                  const int32_t* __restrict__ col, const double* __restrict__ val,
                                                  const int32_t* __restrict__ tptr, int32_t* __restrict__ fill,
                                                  int32_t* __restrict__ tcol, double* __restrict__ tval) {
    int64_t lo, hi;
    block_range(nrows, lo, hi);
    for (int64_t i = lo + threadIdx.x; i < hi; i += kBlock) {
        for (int32_t j = ptr[i]; j < ptr[i + 1]; ++j) {
            const int32_t c = col[j];
            const int32_t pos = tptr[c] + atomicAdd(&fill[c], 1);
            tcol[pos] = (int32_t)i;
            tval[pos] = val[j];
        }
    }
}

// sort each row's entries by column, one wavefront per row: the row is staged in LDS and every
// entry is placed at its rank (columns within a row are distinct); rows longer than the stage
// are sorted by one lane in place (insertion sort)
constexpr int kSortCap = 1024;
__global__ void __launch_bounds__(kWave) k_sort_rows(int64_t nrows, const int32_t* __restrict__ ptr,
                                                     int32_t* __restrict__ col, double* __restrict__ val) {
    __shared__ int32_t sc[kSortCap];
    __shared__ double sv[kSortCap];
    const int lane = threadIdx.x;
    for (int64_t i = blockIdx.x; i < nrows; i += gridDim.x) {
        const int32_t b = ptr[i], e = ptr[i + 1], L = e - b;
        if (L <= 1) continue;
        if (L <= kSortCap) {
            for (int t = lane; t < L; t += kWave) {
                sc[t] = col[b + t];
                sv[t] = val[b + t];
            }
            __syncthreads();
            for (int t = lane; t < L; t += kWave) {
                const int32_t c = sc[t];
                int rank = 0;
                for (int u = 0; u < L; ++u) rank += sc[u] < c ? 1 : 0;
                col[b + rank] = c;
                val[b + rank] = sv[t];
            }
            __syncthreads();
        } else if (lane == 0) {
            for (int32_t j = b + 1; j < e; ++j) {
                const int32_t c = col[j];
                const double v = val[j];
                int32_t q = j - 1;
                while (q >= b && col[q] > c) {
                    col[q + 1] = col[q];
                    val[q + 1] = val[q];
                    --q;
                }
                col[q + 1] = c;
                val[q + 1] = v;
            }
        }
    }
}

// ---------------------------------------------------------------- exclusive scan of int32 counts
__global__ void __launch_bounds__(kBlock) k_scan_blocks(int64_t n, const int32_t* __restrict__ in,
                                                        int64_t* __restrict__ bsum) {
    int64_t lo, hi;
    block_range(n, lo, hi);
    long long acc[1] = {0};
    for (int64_t i = lo + threadIdx.x; i < hi; i += kBlock) acc[0] += in[i];
    block_sum<1>(acc);
    if (threadIdx.x == 0) bsum[blockIdx.x] = acc[0];
}

__global__ void __launch_bounds__(kBlock) k_scan_apply(int64_t n, const int32_t* __restrict__ in,
                                                       const int64_t* __restrict__ bsum, int32_t* __restrict__ out) {
    __shared__ long long base;
    __shared__ long long part[kBlock];
    {
        // the XCD-permuted chunk order: chunks are laid out by logical index, so the base is
        // the sum of the block sums of every block whose logical index is smaller
        const int g = gridDim.x;
        auto logical_of = [g](int bb) { return (g >= kNXcd && g % kNXcd == 0) ? (bb % kNXcd) * (g / kNXcd) + bb / kNXcd : bb; };
        const int me = logical_of(blockIdx.x);
        long long acc[1] = {0};
        for (int bb = threadIdx.x; bb < g; bb += kBlock)
            if (logical_of(bb) < me) acc[0] += bsum[bb];
        block_sum<1>(acc);
        if (threadIdx.x == 0) base = acc[0];
        __syncthreads();
    }
    int64_t lo, hi;
    block_range(n, lo, hi);
    long long run = base;
    for (int64_t start = lo; start < hi; start += kBlock) {
        const int64_t i = start + threadIdx.x;
        const long long v = i < hi ? in[i] : 0;
        part[threadIdx.x] = v;
        __syncthreads();
        // Hillis-Steele inclusive scan in LDS
        for (int off = 1; off < kBlock; off <<= 1) {
            const long long t = threadIdx.x >= off ? part[threadIdx.x - off] : 0;
            __syncthreads();
            part[threadIdx.x] += t;
            __syncthreads();
        }
        if (i < hi) out[i] = (int32_t)(run + part[threadIdx.x] - v);
        run += part[kBlock - 1];
        __syncthreads();
    }
    if (hi == n && lo < hi && threadIdx.x == 0) out[n] = (int32_t)run;
    if (n == 0 && blockIdx.x == 0 && threadIdx.x == 0) out[0] = 0;
}

}  // namespace

extern "C" {

int ftmi_amgs_ldu_to_csr(int64_t n, const int32_t* cptr, const int32_t* he_slot, const int32_t* slot_diag,
                         const int32_t* indptr, const double* diag, const double* off, double* data, void* stream) {
    k_ldu_to_csr<<<grid_for(n), kBlock, 0, (hipStream_t)stream>>>(n, cptr, he_slot, slot_diag, indptr, diag, off,
                                                                   data);
    FTMI_LAUNCH_CHECK("ftmi_amgs_ldu_to_csr");
    return 0;
}

int ftmi_amgs_diag(int64_t n, int64_t nnz, const int32_t* ptr, const int32_t* col, const double* val, double* diag,
                   double* dinv, void* stream) {
    hipStream_t s = (hipStream_t)stream;
    const int tpr = rows_tpr(n, nnz);
    ROWS_TPR_LAUNCH(tpr, k_diag, rows_grid(n, tpr), s, n, ptr, col, val, diag, dinv);
    FTMI_LAUNCH_CHECK("ftmi_amgs_diag");
    return 0;
}

int ftmi_amgs_strength(int64_t n, int64_t nnz, const int32_t* ptr, const int32_t* col, const double* val,
                       const double* diag, double theta, uint8_t* mask, void* stream) {
    hipStream_t s = (hipStream_t)stream;
    const int tpr = rows_tpr(n, nnz);
    ROWS_TPR_LAUNCH(tpr, k_strength, rows_grid(n, tpr), s, n, ptr, col, val, diag, theta, mask,
                    (const uint8_t*)nullptr, (int32_t*)nullptr, 0);
    FTMI_LAUNCH_CHECK("ftmi_amgs_strength");
    return 0;
}

int ftmi_amgs_strength_cmp(int64_t n, int64_t nnz, const int32_t* ptr, const int32_t* col, const double* val,
                           const double* diag, double theta, uint8_t* mask, const uint8_t* old, int32_t* flag,
                           int gen, void* stream) {
    if (old == nullptr || flag == nullptr) {
        set_error_msg("ftmi_amgs_strength_cmp: old mask and flag are required");
        return -1;
    }
    hipStream_t s = (hipStream_t)stream;
    const int tpr = rows_tpr(n, nnz);
    ROWS_TPR_LAUNCH(tpr, k_strength, rows_grid(n, tpr), s, n, ptr, col, val, diag, theta, mask, old, flag, gen);
    FTMI_LAUNCH_CHECK("ftmi_amgs_strength_cmp");
    return 0;
}

int ftmi_differs_u64(int64_t n, const void* a, const void* b, int32_t* flag, int gen, void* stream) {
    if (n <= 0) return 0;
    if (a == nullptr || b == nullptr || flag == nullptr) {
        set_error_msg("ftmi_differs_u64: null buffer");
        return -1;
    }
    k_differs_u64<<<grid_for(n), kBlock, 0, (hipStream_t)stream>>>(n, (const uint64_t*)a, (const uint64_t*)b, flag,
                                                                     gen);
    FTMI_LAUNCH_CHECK("ftmi_differs_u64");
    return 0;
}

int ftmi_amgs_strong_count(int64_t n, const int32_t* ptr, const uint8_t* mask, int32_t* cnt, void* stream) {
    if (n > 0) k_strong_count<<<grid_for(n), kBlock, 0, (hipStream_t)stream>>>(n, ptr, mask, cnt);
    FTMI_LAUNCH_CHECK("ftmi_amgs_strong_count");
    return 0;
}

int ftmi_amgs_strong_fill(int64_t n, const int32_t* ptr, const int32_t* col, const uint8_t* mask,
                          const int32_t* s_ptr, int32_t* s_col, void* stream) {
    if (n > 0) k_strong_fill<<<grid_for(n), kBlock, 0, (hipStream_t)stream>>>(n, ptr, col, mask, s_ptr, s_col);
    FTMI_LAUNCH_CHECK("ftmi_amgs_strong_fill");
    return 0;
}

int ftmi_amgs_tentative(int64_t n, const int32_t* agg, const int32_t* agg_size, int32_t* t_ptr, double* t_val,
                        void* stream) {
    k_tentative<<<grid_for(n + 1), kBlock, 0, (hipStream_t)stream>>>(n, agg, agg_size, t_ptr, t_val);
    FTMI_LAUNCH_CHECK("ftmi_amgs_tentative");
    return 0;
}

// rho(D^-1 A) by `iters` power iterations from x (overwritten); y: 2 n doubles of work;
// partials: >= 2 * 1024 * 2 doubles; out[0] = lam (device).
int ftmi_amgs_power(int64_t n, int64_t nnz, const int32_t* ptr, const int32_t* col, const double* val,
                    const double* dinv, double* x, double* y, int iters, double* partials, double* out, void* stream) {
    if (n <= 0 || iters <= 0) return 0;
    hipStream_t s = (hipStream_t)stream;
    // rows of up to 24 entries keep one lane, so the estimate is the column-order (reference)
    // sum bit for bit where the hierarchy is parity-checked; longer rows take rows_tpr lanes
    const int tpr = (double)nnz <= 24.0 * (double)n ? 1 : rows_tpr(n, nnz);
    int64_t g64 = (n * tpr + kBlock - 1) / kBlock;
    const int g = (int)(g64 < 1 ? 1 : (g64 > kRedGrid ? kRedGrid : g64));
    double* part[2] = {partials, partials + 2 * kRedGrid};
    double* yb[2] = {y, y + n};
    for (int it = 0; it < iters; ++it) {
        const int a = it & 1, c = a ^ 1;
        if (it == 0)
            ROWS_TPR_LAUNCH(tpr, k_pow_first, g, s, n, ptr, col, val, dinv, yb[c], x, yb[a], part[c], g, part[a]);
        else
            ROWS_TPR_LAUNCH(tpr, k_pow_step, g, s, n, ptr, col, val, dinv, yb[c], x, yb[a], part[c], g, part[a]);
    }
    k_pow_fin<<<1, kBlock, 0, s>>>(part[(iters - 1) & 1], g, &out[0]);
    FTMI_LAUNCH_CHECK("ftmi_amgs_power");
    return 0;
}

// Row counts (fill == 0) or entries (fill == 1) of C = diag(xscale) X Y (C has ncols columns);
// with agg != NULL, C = T - (omega_scale / *omega_num) diag(xscale) X Y with T = (agg, tval).
// overflow (4 int) records rows whose distinct columns exceed the LDS hash capacity (count,
// max, bound, row). Small products (kDenseCols columns, kDenseWork rows x columns) use the
// dense-accumulator kernel.
int ftmi_amgs_spgemm(int fill, int64_t nrows, int64_t ncols, const int32_t* x_ptr, const int32_t* x_col,
                     const double* x_val,
                     const int32_t* y_ptr, const int32_t* y_col, const double* y_val, const double* xscale,
                     const int32_t* agg, const double* tval, const double* omega_num, double omega_scale,
                     int32_t* rowcnt, const int32_t* c_ptr, int32_t* c_col, double* c_val, int* overflow,
                     int32_t* long_rows, int32_t* n_long, void* stream) {
    if (nrows <= 0) return 0;
    if (!rowcnt || !long_rows || !n_long || (fill && (!c_ptr || !c_col || !c_val))) {
        set_error_msg("ftmi_amgs_spgemm: missing work or output buffer");
        return -1;
    }
    hipStream_t s = (hipStream_t)stream;
    CsrV X{x_ptr, x_col, x_val}, Y{y_ptr, y_col, y_val};
    const int g = grid_for(nrows);
    const int gl = nrows < kMaxGrid ? (int)nrows : kMaxGrid;
    if (ncols <= kDenseCols && nrows * ncols <= kDenseWork) {
        const int gd = nrows < 2 * kMaxGrid ? (int)nrows : 2 * kMaxGrid;
        const int nc = (int)ncols;
        const size_t words = (size_t)((nc + 31) / 32) * 4;
        if (fill)
            k_spgemm_dense<true><<<gd, kWave, (size_t)nc * 8 + words, s>>>(nrows, nc, X, Y, xscale, agg, tval,
                                                                         omega_num, omega_scale, rowcnt, c_ptr,
                                                                         c_col, c_val);
        else
            k_spgemm_dense<false><<<gd, kWave, words, s>>>(nrows, nc, X, Y, xscale, agg, tval, omega_num,
                                                         omega_scale, rowcnt, c_ptr, c_col, c_val);
    } else if (fill) {
        k_spgemm_merge<true><<<g, kBlock, 0, s>>>(nrows, X, Y, xscale, agg, tval, omega_num, omega_scale, rowcnt,
                                                  c_ptr, c_col, c_val, long_rows, n_long);
        k_spgemm<true><<<gl, kWave, 0, s>>>(long_rows, n_long, X, Y, xscale, agg, tval, omega_num, omega_scale,
                                            rowcnt, c_ptr, c_col, c_val, overflow);
    } else {
        k_spgemm_merge<false><<<g, kBlock, 0, s>>>(nrows, X, Y, xscale, agg, tval, omega_num, omega_scale, rowcnt,
                                                   c_ptr, c_col, c_val, long_rows, n_long);
        k_spgemm<false><<<gl, kWave, 0, s>>>(long_rows, n_long, X, Y, xscale, agg, tval, omega_num, omega_scale,
                                             rowcnt, c_ptr, c_col, c_val, overflow);
    }
    FTMI_LAUNCH_CHECK("ftmi_amgs_spgemm");
    return 0;
}

// out[0..n] = exclusive prefix sums of in[0..n-1] (out[n] = total); bsum: >= 2048 int64.
int ftmi_scan_i32(int64_t n, const int32_t* in, int32_t* out, int64_t* bsum, void* stream) {
    hipStream_t s = (hipStream_t)stream;
    const int g = grid_part(n > 0 ? n : 1);
    k_scan_blocks<<<g, kBlock, 0, s>>>(n, in, bsum);
    k_scan_apply<<<g, kBlock, 0, s>>>(n, in, bsum, out);
    FTMI_LAUNCH_CHECK("ftmi_scan_i32");
    return 0;
}

// R = P^T for P (nrows x ncols) in CSR; r_ptr must hold the exclusive scan of the column
// counts (ftmi_amgs_count_cols + ftmi_scan_i32); fill: ncols int32 zeros (work).
int ftmi_amgs_count_cols(int64_t nnz, const int32_t* col, int32_t* cnt, void* stream) {
    if (nnz > 0) k_count_cols<<<grid_for(nnz), kBlock, 0, (hipStream_t)stream>>>(nnz, col, cnt);
    FTMI_LAUNCH_CHECK("ftmi_amgs_count_cols");
    return 0;
}

int ftmi_amgs_transpose_fill(int64_t nrows, int64_t ncols, const int32_t* ptr, const int32_t* col, const double* val,
                             const int32_t* r_ptr, int32_t* fill, int32_t* r_col, double* r_val, void* stream) {
    hipStream_t s = (hipStream_t)stream;
    k_place<<<grid_for(nrows), kBlock, 0, s>>>(nrows, ptr, col, val, r_ptr, fill, r_col, r_val);
    const int64_t gs = ncols < 65536 ? (ncols > 0 ? ncols : 1) : 65536;
    k_sort_rows<<<(int)gs, kWave, 0, s>>>(ncols, r_ptr, r_col, r_val);
    FTMI_LAUNCH_CHECK("ftmi_amgs_transpose_fill");
    return 0;
}

}  // extern "C"
