// Shared device helpers of libftmi (gfx950, wave64).
//
// Launch shape: every cell/face kernel runs 256-thread workgroups over a capped grid;
// workgroup b handles one contiguous chunk of rows, and the chunks are assigned so that
// the workgroups one XCD receives (the dispatcher deals workgroups round-robin over the
// 8 XCDs) cover one contiguous range of rows: neighbour gathers then stay in that XCD's L2.
//
// Reductions are deterministic: each workgroup reduces its rows with wave shuffles and
// LDS in a fixed tree and writes its partial sums; a one-workgroup finalize kernel adds the
// partials in a fixed order and runs the scalar bookkeeping. No atomics anywhere.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include <cmath>

namespace ftmi {

constexpr int kBlock = 256;
constexpr int kWave = 64;
constexpr int kMaxGrid = 2048;  // 8 per CU on 256 CUs
constexpr int kNXcd = 8;
constexpr int kMaxRed = 8;  // values reduced together by one kernel

// Per-solve device workspace for reductions: partials[kMaxGrid * kMaxRed] (the counter is
// unused and kept for the layout of the C struct).
struct RedBuf {
    double* partials;
    unsigned* counter;
};

// Workgroup cap of the streaming kernels (kMaxGrid by default; the autotuner may change it
// through ftmi_set_grid_cap, up to 65536).
int grid_cap();

inline int grid_for(int64_t n) {
    int64_t g = (n + kBlock - 1) / kBlock;
    if (g < 1) g = 1;
    const int cap = grid_cap();
    if (g > cap) g = cap;
    return (int)g;
}

// Grid of a kernel that writes per-workgroup values into a caller's buffer sized for kMaxGrid
// workgroups (RedBuf partials: kMaxGrid * kMaxRed doubles; scan block sums: kMaxGrid): the
// streaming grid, but never more than kMaxGrid workgroups whatever the cap.
inline int grid_part(int64_t n) {
    const int g = grid_for(n);
    return g < kMaxGrid ? g : kMaxGrid;
}

// Contiguous row range [lo, hi) of this workgroup with the XCD-aware block permutation.
__device__ inline void block_range(int64_t n, int64_t& lo, int64_t& hi) {
    const int g = gridDim.x;
    const int b = blockIdx.x;
    int logical = b;
    if (g >= kNXcd && g % kNXcd == 0) logical = (b % kNXcd) * (g / kNXcd) + b / kNXcd;
    const int64_t chunk = ((n + g - 1) / g + kWave - 1) / kWave * kWave;
    lo = (int64_t)logical * chunk;
    hi = lo + chunk;
    if (hi > n) hi = n;
    if (lo > n) lo = n;
}

template <typename T>
__device__ inline T wave_sum(T v) {
#pragma unroll
    for (int off = kWave / 2; off > 0; off >>= 1) v += __shfl_xor(v, off, kWave);
    return v;
}

// Sum M values over the workgroup (fixed order); result valid in thread 0.
template <int M, typename T>
__device__ inline void block_sum(T (&v)[M]) {
    __shared__ T lds[kBlock / kWave][M > 0 ? M : 1];
    const int lane = threadIdx.x % kWave, w = threadIdx.x / kWave;
#pragma unroll
    for (int j = 0; j < M; ++j) v[j] = wave_sum(v[j]);
    if (lane == 0) {
#pragma unroll
        for (int j = 0; j < M; ++j) lds[w][j] = v[j];
    }
    __syncthreads();
    if (threadIdx.x == 0) {
#pragma unroll
        for (int j = 0; j < M; ++j) {
            T s = lds[0][j];
            for (int k = 1; k < kBlock / kWave; ++k) s += lds[k][j];
            v[j] = s;
        }
    }
    __syncthreads();
}

// Producer side of a grid reduction: the workgroup's sums go to partials[blockIdx.x * M + j].
// No fence and no atomic: the consumer is a separate one-workgroup kernel (finalize_sum), so
// the kernel boundary orders the writes. (A per-workgroup agent-scope fence would write back
// the XCD's L2 every time, which measured ~8x slower than the reduction-free kernel.)
template <int M, typename T>
__device__ inline void block_partials(T (&v)[M], double* partials) {
    block_sum<M>(v);
    if (threadIdx.x == 0) {
#pragma unroll
        for (int j = 0; j < M; ++j) partials[(int64_t)blockIdx.x * M + j] = (double)v[j];
    }
}

// Consumer side, run by the single workgroup of a finalize kernel: sums the `nparts` partials
// in a fixed order (strided per thread, then a fixed tree); true in thread 0, totals in `tot`.
template <int M>
__device__ inline bool finalize_sum(const double* partials, int nparts, double (&tot)[M]) {
    double acc[M];
#pragma unroll
    for (int j = 0; j < M; ++j) acc[j] = 0.0;
    for (int b = threadIdx.x; b < nparts; b += kBlock) {
#pragma unroll
        for (int j = 0; j < M; ++j) acc[j] += partials[(int64_t)b * M + j];
    }
    block_sum<M>(acc);
    if (threadIdx.x == 0) {
#pragma unroll
        for (int j = 0; j < M; ++j) tot[j] = acc[j];
    }
    return threadIdx.x == 0;
}

// Consumer-side total computed by every workgroup of the consuming kernel (all threads get
// it): the producer ran at most kRedGrid workgroups, so each consumer workgroup re-adds the
// partials in the same fixed order and every workgroup sees bit-identical totals.
constexpr int kRedGrid = 1024;
inline int grid_red(int64_t n) {
    const int g = grid_for(n);
    return g < kRedGrid ? g : kRedGrid;
}

template <int M>
__device__ inline void consume_sum(const double* partials, int nparts, double (&tot)[M]) {
    __shared__ double bc[M];
    double t[M];
    if (finalize_sum<M>(partials, nparts, t)) {
#pragma unroll
        for (int j = 0; j < M; ++j) bc[j] = t[j];
    }
    __syncthreads();
#pragma unroll
    for (int j = 0; j < M; ++j) tot[j] = bc[j];
    __syncthreads();
}

// Generic finalize: out[j] = scale * total[j].
template <int M>
__global__ void __launch_bounds__(kBlock) fin_store(const double* partials, int nparts, double* out, double scale) {
    double tot[M];
    if (finalize_sum<M>(partials, nparts, tot)) {
#pragma unroll
        for (int j = 0; j < M; ++j) out[j] = scale * tot[j];
    }
}

}  // namespace ftmi

// Error bookkeeping shared by every entry point (thread-local last error text).
namespace ftmi {
void set_error(const char* where, hipError_t e);
void set_error_msg(const char* msg);
}  // namespace ftmi

#define FTMI_LAUNCH_CHECK(where)                          \
    do {                                                  \
        hipError_t _e = hipGetLastError();                \
        if (_e != hipSuccess) {                           \
            ftmi::set_error(where, _e);                   \
            return -(int)_e;                              \
        }                                                 \
    } while (0)
