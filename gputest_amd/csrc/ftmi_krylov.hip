// Device-resident Krylov solvers and the smoothed-aggregation V-cycle.
//
// * Jacobi-right-preconditioned BiCGSTAB, the algorithm and stopping rules of
//   flowtrace/linalg/krylov.py:68-121 (and of the reference's C++ port,
//   flowtrace/native/fvm_kernels.cpp:146-234): stop when ||r||/||r0|| < tol (or ||r|| <= 1e-14 ||b||),
//   early exit on ||s||, breakdown keeps the last finite iterate.
// * PCG with null-space projection, flowtrace/linalg/krylov.py:25-65.
// * V(2,2) weighted-Jacobi cycle, flowtrace/linalg/amg.py:130-146.
//
// Every scalar of the iteration (rho, alpha, omega, norms, flags) lives in device memory and
// is computed by the last workgroup of the reducing kernel, so an iteration is a fixed
// sequence of launches with no host round trip. Each kernel first reads the device `done`
// flag (set only by an earlier kernel) and returns at once when the solve has finished.
// The host checks `done` one iteration behind, through an asynchronous copy into pinned
// memory, so the GPU never waits for the host.
//
// The BiCGSTAB iterate lives in two buffers (the caller's x and a work buffer): the update
// x + alpha phat + omega shat is written to the buffer that is not current and becomes
// current only when ||r|| is finite, which is the reference's "keep the last finite
// iterate" rule without an extra pass.
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <vector>

#include "../../include/ftmi.h"
#include "ftmi_common.h"
#include "ftmi_ops.h"

using namespace ftmi;

namespace {

// ------------------------------------------------------------------ device state block
enum { S_RHO, S_ALPHA, S_OMEGA, S_RNORM, S_BNORM, S_BABS, S_RHONEW, S_RES, S_RZ, S_BETA, S_ZMEAN, S_BMEAN, S_NS = 16 };
enum { I_DONE, I_CONV, I_ITERS, I_CUR, I_FIXUP, I_NI = 8 };

struct State {  // 192 bytes, copied to the host as one block
    double sc[S_NS];
    int st[I_NI];
    int pad[8];
};

__global__ void k_state_reset(State* s, unsigned* counter) {
    const int t = threadIdx.x;
    if (t < S_NS) s->sc[t] = 0.0;
    if (t < I_NI) s->st[t] = 0;
    if (t == 0) *counter = 0u;
}

// ------------------------------------------------------------------ host-side flag reader
struct Pinned {
    State* h = nullptr;  // two slots
    hipEvent_t ev[2] = {nullptr, nullptr};
    bool ok = false;
    Pinned() {
        if (hipHostMalloc((void**)&h, 2 * sizeof(State), hipHostMallocDefault) != hipSuccess) return;
        if (hipEventCreateWithFlags(&ev[0], hipEventDisableTiming) != hipSuccess) return;
        if (hipEventCreateWithFlags(&ev[1], hipEventDisableTiming) != hipSuccess) return;
        ok = true;
    }
};
static thread_local Pinned* g_pinned = nullptr;
static Pinned* pinned() {
    if (!g_pinned) g_pinned = new Pinned();
    return g_pinned->ok ? g_pinned : nullptr;
}

static int post_state(Pinned* pn, int slot, const State* d, hipStream_t s) {
    if (hipMemcpyAsync(&pn->h[slot], d, sizeof(State), hipMemcpyDeviceToHost, s) != hipSuccess) return -1;
    if (hipEventRecord(pn->ev[slot], s) != hipSuccess) return -1;
    return 0;
}
static int wait_state(Pinned* pn, int slot) { return hipEventSynchronize(pn->ev[slot]) == hipSuccess ? 0 : -1; }

// ------------------------------------------------------------------ BiCGSTAB kernels
template <class Op>
__global__ void __launch_bounds__(kBlock) bi_init(int64_t n, Op A, const double* __restrict__ b,
                                                  const double* __restrict__ x, double* __restrict__ r,
                                                  double* __restrict__ rh, double* __restrict__ p,
                                                  double* __restrict__ v, State* S, RedBuf rb, double tol, int rel) {
    int64_t lo, hi;
    block_range(n, lo, hi);
    double acc[2] = {0.0, 0.0};
    for (int64_t i = lo + threadIdx.x; i < hi; i += kBlock) {
        const double ri = b[i] - A.row(i, x);
        r[i] = ri;
        rh[i] = ri;
        p[i] = 0.0;
        v[i] = 0.0;
        acc[0] += ri * ri;
        acc[1] += b[i] * b[i];
    }
    block_partials<2>(acc, rb.partials);
}

__global__ void __launch_bounds__(kBlock) fin_bi_init(const double* partials, int nparts, State* S, double tol,
                                                      int rel) {
    double tot[2];
    if (finalize_sum<2>(partials, nparts, tot)) {
        const double rnorm = sqrt(tot[0]), b_abs = sqrt(tot[1]);
        double bnorm = rel ? rnorm : b_abs;
        if (bnorm == 0.0) bnorm = 1.0;
        double* sc = S->sc;
        int* st = S->st;
        sc[S_RNORM] = rnorm;
        sc[S_BNORM] = bnorm;
        sc[S_BABS] = b_abs;
        sc[S_RHO] = sc[S_ALPHA] = sc[S_OMEGA] = 1.0;
        sc[S_RHONEW] = tot[0];
        sc[S_RES] = rnorm / bnorm;
        st[I_CUR] = 0;
        st[I_FIXUP] = 0;
        if (rnorm / bnorm < tol || rnorm <= 1e-14 * b_abs) {
            st[I_DONE] = 1;
            st[I_CONV] = 1;
            st[I_ITERS] = 0;
        }
    }
}

__global__ void __launch_bounds__(kBlock) bi_p(int64_t n, const double* __restrict__ r, double* __restrict__ p,
                                               const double* __restrict__ v, const double* __restrict__ inv,
                                               double* __restrict__ ph, const State* S) {
    if (S->st[I_DONE]) return;
    const double* sc = S->sc;
    const double beta = (sc[S_RHONEW] / sc[S_RHO]) * (sc[S_ALPHA] / sc[S_OMEGA]);
    const double om = sc[S_OMEGA];
    int64_t lo, hi;
    block_range(n, lo, hi);
    for (int64_t i = lo + threadIdx.x; i < hi; i += kBlock) {
        const double pi = r[i] + beta * (p[i] - om * v[i]);
        p[i] = pi;
        ph[i] = inv[i] * pi;
    }
}

template <class Op>
__global__ void __launch_bounds__(kBlock) bi_v(int64_t n, Op A, const double* __restrict__ ph,
                                               const double* __restrict__ rh, double* __restrict__ v, State* S,
                                               RedBuf rb, int k) {
    if (S->st[I_DONE]) return;
    int64_t lo, hi;
    block_range(n, lo, hi);
    double acc[1] = {0.0};
    for (int64_t i = lo + threadIdx.x; i < hi; i += kBlock) {
        const double vi = A.row(i, ph);
        v[i] = vi;
        acc[0] += rh[i] * vi;
    }
    block_partials<1>(acc, rb.partials);
}

__global__ void __launch_bounds__(kBlock) fin_bi_v(const double* partials, int nparts, State* S, int k) {
    if (S->st[I_DONE]) return;
    double tot[1];
    if (finalize_sum<1>(partials, nparts, tot)) {
        double* sc = S->sc;
        sc[S_RHO] = sc[S_RHONEW];
        if (tot[0] == 0.0) {
            S->st[I_DONE] = 1;
            S->st[I_CONV] = 0;
            S->st[I_ITERS] = k;
            sc[S_RES] = sc[S_RNORM] / sc[S_BNORM];
        } else {
            sc[S_ALPHA] = sc[S_RHO] / tot[0];
        }
    }
}

__global__ void __launch_bounds__(kBlock) bi_s(int64_t n, const double* __restrict__ r, const double* __restrict__ v,
                                               const double* __restrict__ inv, double* __restrict__ s,
                                               double* __restrict__ sh, State* S, RedBuf rb, double tol, int k) {
    if (S->st[I_DONE]) return;
    const double alpha = S->sc[S_ALPHA];
    int64_t lo, hi;
    block_range(n, lo, hi);
    double acc[1] = {0.0};
    for (int64_t i = lo + threadIdx.x; i < hi; i += kBlock) {
        const double si = r[i] - alpha * v[i];
        s[i] = si;
        sh[i] = inv[i] * si;
        acc[0] += si * si;
    }
    block_partials<1>(acc, rb.partials);
}

__global__ void __launch_bounds__(kBlock) fin_bi_s(const double* partials, int nparts, State* S, double tol, int k) {
    if (S->st[I_DONE]) return;
    double tot[1];
    if (finalize_sum<1>(partials, nparts, tot)) {
        double* sc = S->sc;
        const double snorm = sqrt(tot[0]);
        if (snorm / sc[S_BNORM] < tol) {
            S->st[I_DONE] = 1;
            S->st[I_CONV] = 1;
            S->st[I_ITERS] = k;
            S->st[I_FIXUP] = 1;  // x = x_cur + alpha phat, applied after the loop
            sc[S_RES] = snorm / sc[S_BNORM];
        }
    }
}

template <class Op>
__global__ void __launch_bounds__(kBlock) bi_t(int64_t n, Op A, const double* __restrict__ sh,
                                               const double* __restrict__ s, double* __restrict__ t, State* S,
                                               RedBuf rb) {
    if (S->st[I_DONE]) return;
    int64_t lo, hi;
    block_range(n, lo, hi);
    double acc[2] = {0.0, 0.0};
    for (int64_t i = lo + threadIdx.x; i < hi; i += kBlock) {
        const double ti = A.row(i, sh);
        t[i] = ti;
        acc[0] += ti * ti;
        acc[1] += ti * s[i];
    }
    block_partials<2>(acc, rb.partials);
}

__global__ void __launch_bounds__(kBlock) fin_bi_t(const double* partials, int nparts, State* S) {
    if (S->st[I_DONE]) return;
    double tot[2];
    if (finalize_sum<2>(partials, nparts, tot)) S->sc[S_OMEGA] = tot[0] > 0.0 ? tot[1] / tot[0] : 0.0;
}

__global__ void __launch_bounds__(kBlock) bi_r(int64_t n, const double* __restrict__ s, const double* __restrict__ t,
                                               const double* __restrict__ ph, const double* __restrict__ sh,
                                               const double* __restrict__ rh, double* __restrict__ r,
                                               double* __restrict__ x0, double* __restrict__ x1, State* S, RedBuf rb,
                                               double tol, int k, int maxiter) {
    if (S->st[I_DONE]) return;
    const double alpha = S->sc[S_ALPHA], om = S->sc[S_OMEGA];
    const int cur = S->st[I_CUR];
    const double* xc = cur ? x1 : x0;
    double* xo = cur ? x0 : x1;
    int64_t lo, hi;
    block_range(n, lo, hi);
    double acc[2] = {0.0, 0.0};
    for (int64_t i = lo + threadIdx.x; i < hi; i += kBlock) {
        const double ri = s[i] - om * t[i];
        r[i] = ri;
        xo[i] = xc[i] + alpha * ph[i] + om * sh[i];
        acc[0] += ri * ri;
        acc[1] += rh[i] * ri;
    }
    block_partials<2>(acc, rb.partials);
}

__global__ void __launch_bounds__(kBlock) fin_bi_r(const double* partials, int nparts, State* S, double tol, int k,
                                                   int maxiter) {
    if (S->st[I_DONE]) return;
    double tot[2];
    if (finalize_sum<2>(partials, nparts, tot)) {
        double* sc = S->sc;
        int* st = S->st;
        const int cur = st[I_CUR];
        const double om = sc[S_OMEGA];
        const double rnorm = sqrt(tot[0]);
        if (!isfinite(rnorm)) {  // breakdown: the current iterate stays
            st[I_DONE] = 1;
            st[I_CONV] = 0;
            st[I_ITERS] = k;
            sc[S_RES] = __builtin_nan("");
            return;
        }
        st[I_CUR] = cur ^ 1;
        sc[S_RNORM] = rnorm;
        sc[S_RES] = rnorm / sc[S_BNORM];
        if (rnorm / sc[S_BNORM] < tol || rnorm <= 1e-14 * sc[S_BABS]) {
            st[I_DONE] = 1;
            st[I_CONV] = 1;
            st[I_ITERS] = k;
        } else if (om == 0.0) {
            st[I_DONE] = 1;
            st[I_CONV] = 0;
            st[I_ITERS] = k;
        } else {
            sc[S_RHONEW] = tot[1];
            if (tot[1] == 0.0 && k < maxiter) {  // the next iteration breaks at its first step
                st[I_DONE] = 1;
                st[I_CONV] = 0;
                st[I_ITERS] = k + 1;
            }
        }
    }
}

// x0 = x_cur + alpha phat (early exit on ||s||), or x0 = x1 when x1 is current.
__global__ void __launch_bounds__(kBlock) bi_final(int64_t n, double* __restrict__ x0, const double* __restrict__ x1,
                                                   const double* __restrict__ ph, double alpha, int cur, int fixup) {
    int64_t lo, hi;
    block_range(n, lo, hi);
    for (int64_t i = lo + threadIdx.x; i < hi; i += kBlock) {
        const double xc = cur ? x1[i] : x0[i];
        x0[i] = fixup ? xc + alpha * ph[i] : xc;
    }
}

struct BiWork {
    double *r, *rh, *p, *v, *ph, *s, *sh, *t, *x1;
    State* S;
    RedBuf rb;
};

static BiWork carve_bi(int64_t n, void* work) {
    char* w = (char*)work;
    BiWork W;
    double* d = (double*)w;
    W.r = d;
    W.rh = d + n;
    W.p = d + 2 * n;
    W.v = d + 3 * n;
    W.ph = d + 4 * n;
    W.s = d + 5 * n;
    W.sh = d + 6 * n;
    W.t = d + 7 * n;
    W.x1 = d + 8 * n;
    char* tail = w + ((9 * n * sizeof(double) + 255) / 256) * 256;
    W.S = (State*)tail;
    W.rb.counter = (unsigned*)(tail + 256);
    W.rb.partials = (double*)(tail + 512);
    return W;
}

}  // namespace

extern "C" int64_t ftmi_bicgstab_work_bytes(int64_t n) {
    return ((9 * n * (int64_t)sizeof(double) + 255) / 256) * 256 + 512 + (int64_t)kMaxGrid * kMaxRed * 8;
}

namespace ftmi {

// Runs the whole solve; returns 1 if converged, 0 if not, <0 on a HIP error.
template <class Op>
int bicgstab_run(int64_t n, Op A, const double* b, const double* inv, double* x, double tol, int64_t maxiter,
                 int rel, void* work, int64_t* iters, double* res, hipStream_t st) {
    Pinned* pn = pinned();
    if (!pn) {
        set_error_msg("bicgstab: pinned host memory unavailable");
        return -1;
    }
    BiWork W = carve_bi(n, work);
    const int g = grid_for(n);
    k_state_reset<<<1, 64, 0, st>>>(W.S, W.rb.counter);
    bi_init<Op><<<g, kBlock, 0, st>>>(n, A, b, x, W.r, W.rh, W.p, W.v, W.S, W.rb, tol, rel);
    fin_bi_init<<<1, kBlock, 0, st>>>(W.rb.partials, g, W.S, tol, rel);
    FTMI_LAUNCH_CHECK("bicgstab init");
    if (post_state(pn, 0, W.S, st)) return -1;
    // slot `last` holds the most recently posted state; the loop waits one iteration behind
    int last = 0;
    for (int64_t k = 1; k <= maxiter; ++k) {
        bi_p<<<g, kBlock, 0, st>>>(n, W.r, W.p, W.v, inv, W.ph, W.S);
        bi_v<Op><<<g, kBlock, 0, st>>>(n, A, W.ph, W.rh, W.v, W.S, W.rb, (int)k);
        fin_bi_v<<<1, kBlock, 0, st>>>(W.rb.partials, g, W.S, (int)k);
        bi_s<<<g, kBlock, 0, st>>>(n, W.r, W.v, inv, W.s, W.sh, W.S, W.rb, tol, (int)k);
        fin_bi_s<<<1, kBlock, 0, st>>>(W.rb.partials, g, W.S, tol, (int)k);
        bi_t<Op><<<g, kBlock, 0, st>>>(n, A, W.sh, W.s, W.t, W.S, W.rb);
        fin_bi_t<<<1, kBlock, 0, st>>>(W.rb.partials, g, W.S);
        bi_r<<<g, kBlock, 0, st>>>(n, W.s, W.t, W.ph, W.sh, W.rh, W.r, x, W.x1, W.S, W.rb, tol, (int)k, (int)maxiter);
        fin_bi_r<<<1, kBlock, 0, st>>>(W.rb.partials, g, W.S, tol, (int)k, (int)maxiter);
        FTMI_LAUNCH_CHECK("bicgstab iteration");
        if (post_state(pn, last ^ 1, W.S, st)) return -1;
        if (wait_state(pn, last)) return -1;  // the previous iteration's flags
        const bool prev_done = pn->h[last].st[I_DONE] != 0;
        last ^= 1;
        if (prev_done) break;
    }
    // every posted state is final once the solve is done; the last one is the latest
    if (wait_state(pn, last)) return -1;
    const int slot = last;
    const State& H = pn->h[slot];
    int conv;
    if (H.st[I_DONE]) {
        *iters = H.st[I_ITERS];
        *res = H.sc[S_RES];
        conv = H.st[I_CONV];
    } else {
        *iters = maxiter;
        *res = H.sc[S_RES];
        conv = 0;
    }
    if (H.st[I_CUR] || H.st[I_FIXUP])
        bi_final<<<g, kBlock, 0, st>>>(n, x, W.x1, W.ph, H.sc[S_ALPHA], H.st[I_CUR], H.st[I_FIXUP]);
    FTMI_LAUNCH_CHECK("bicgstab final");
    return conv;
}

template int bicgstab_run<CsrOp<double>>(int64_t, CsrOp<double>, const double*, const double*, double*, double,
                                         int64_t, int, void*, int64_t*, double*, hipStream_t);
template int bicgstab_run<LduOp<double>>(int64_t, LduOp<double>, const double*, const double*, double*, double,
                                         int64_t, int, void*, int64_t*, double*, hipStream_t);

}  // namespace ftmi

extern "C" {

int ftmi_bicgstab_jacobi(int64_t n, const int32_t* indptr, const int32_t* indices, const double* data,
                         const double* b, const double* inv_diag, double* x, double tol, int64_t maxiter,
                         int relative_to_initial, void* work, int64_t* iters, double* res, void* stream) {
    if (n <= 0) {
        *iters = 0;
        *res = 0.0;
        return 1;
    }
    CsrOp<double> A{indptr, indices, data};
    return bicgstab_run(n, A, b, inv_diag, x, tol, maxiter, relative_to_initial, work, iters, res,
                        (hipStream_t)stream);
}

int ftmi_bicgstab_ldu(int64_t n, const double* diag, const int32_t* ptr, const int32_t* nbr, const double* off,
                      const double* b, const double* inv_diag, double* x, double tol, int64_t maxiter,
                      int relative_to_initial, void* work, int64_t* iters, double* res, void* stream) {
    if (n <= 0) {
        *iters = 0;
        *res = 0.0;
        return 1;
    }
    LduOp<double> A{diag, ptr, nbr, off};
    return bicgstab_run(n, A, b, inv_diag, x, tol, maxiter, relative_to_initial, work, iters, res,
                        (hipStream_t)stream);
}

}  // extern "C"

// ====================================================================== AMG V-cycle
namespace {

struct Level {
    int64_t n;
    CsrOp<double> A, P, R;
    int tpr_a, tpr_p, tpr_r;  // threads per row of the level's CSR kernels
    int64_t nnz_p;
    const double* Dinv;
    double *b, *x, *y, *r;  // level vectors (level 0: b is the caller's rhs, x its output)
};

struct Amg {
    int nlev = 0;
    Level lv[24];
    int64_t nc = 0;
    const double* cinv = nullptr;
    double *cb = nullptr, *cx = nullptr;
    double omega = 0.8;
    int sweeps = 2;
};

// Rows of the coarse Galerkin levels grow long (hundreds of entries on the last levels), so
// each CSR kernel of the cycle gives TPR lanes to a row and reduces them with width-TPR
// shuffles; TPR = 1 keeps the plain column-order sum of the fine levels.
inline int pick_tpr(int64_t n, int64_t nnz) {
    const double avg = n > 0 ? (double)nnz / (double)n : 0.0;
    if (avg <= 8.0) return 1;
    if (avg <= 24.0) return 4;
    if (avg <= 96.0) return 16;
    return 64;
}

template <int TPR>
struct RowGroups {  // row r handled by lanes [g*TPR, g*TPR + TPR) of the grid
    int64_t first, stride;
    int lane;
    __device__ RowGroups() {
        const int g = gridDim.x, b = blockIdx.x;
        int logical = b;
        if (g >= kNXcd && g % kNXcd == 0) logical = (b % kNXcd) * (g / kNXcd) + b / kNXcd;
        first = (int64_t)logical * (kBlock / TPR) + threadIdx.x / TPR;
        stride = (int64_t)g * (kBlock / TPR);
        lane = threadIdx.x % TPR;
    }
};

template <int TPR>
__device__ inline double seg_sum(double v) {
#pragma unroll
    for (int off = TPR / 2; off > 0; off >>= 1) v += __shfl_xor(v, off, TPR);
    return v;
}

template <int TPR>
__device__ inline double row_dot(const CsrOp<double>& A, int64_t i, const double* __restrict__ x, int lane) {
    double s = 0.0;
    const int32_t e = A.ptr[i + 1];
    for (int32_t j = A.ptr[i] + lane; j < e; j += TPR) s += A.val[j] * x[A.col[j]];
    return seg_sum<TPR>(s);
}

inline int grid_rows(int64_t n, int tpr) { return grid_for(n * tpr); }

// x = w D b + w D (b - A (w D b)): the first two pre-smoothing sweeps from x = 0, fused.
template <int TPR>
__global__ void __launch_bounds__(kBlock) vc_pre2(int64_t n, CsrOp<double> A, const double* __restrict__ Dinv,
                                                  const double* __restrict__ b, double* __restrict__ x, double w,
                                                  const int* done) {
    if (done && *done) return;
    RowGroups<TPR> G;
    for (int64_t i = G.first; i < n; i += G.stride) {
        double ax = 0.0;
        const int32_t e = A.ptr[i + 1];
        for (int32_t j = A.ptr[i] + G.lane; j < e; j += TPR) {
            const int32_t c = A.col[j];
            ax += A.val[j] * ((w * Dinv[c]) * b[c]);
        }
        ax = seg_sum<TPR>(ax);
        if (G.lane == 0) {
            const double wd = w * Dinv[i];
            x[i] = wd * b[i] + wd * (b[i] - ax);
        }
    }
}

// x = w D b (first sweep only; used when sweeps == 1)
__global__ void __launch_bounds__(kBlock) vc_pre1(int64_t n, const double* __restrict__ Dinv,
                                                  const double* __restrict__ b, double* __restrict__ x, double w,
                                                  const int* done) {
    if (done && *done) return;
    int64_t lo, hi;
    block_range(n, lo, hi);
    for (int64_t i = lo + threadIdx.x; i < hi; i += kBlock) x[i] = (w * Dinv[i]) * b[i];
}

// y = x + w D (b - A x)   (optionally also the workgroup partial sums of y)
template <int TPR, bool SUM>
__global__ void __launch_bounds__(kBlock) vc_jacobi(int64_t n, CsrOp<double> A, const double* __restrict__ Dinv,
                                                    const double* __restrict__ b, const double* __restrict__ x,
                                                    double* __restrict__ y, double w, const int* done, RedBuf rb,
                                                    double* mean_out) {
    (void)mean_out;
    if (done && *done) return;
    RowGroups<TPR> G;
    double acc[1] = {0.0};
    for (int64_t i = G.first; i < n; i += G.stride) {
        const double ax = row_dot<TPR>(A, i, x, G.lane);
        if (G.lane == 0) {
            const double yi = x[i] + (w * Dinv[i]) * (b[i] - ax);
            y[i] = yi;
            if (SUM) acc[0] += yi;
        }
    }
    if (SUM) block_partials<1>(acc, rb.partials);
}

template <int TPR>
__global__ void __launch_bounds__(kBlock) vc_resid(int64_t n, CsrOp<double> A, const double* __restrict__ b,
                                                   const double* __restrict__ x, double* __restrict__ r,
                                                   const int* done) {
    if (done && *done) return;
    RowGroups<TPR> G;
    for (int64_t i = G.first; i < n; i += G.stride) {
        const double ax = row_dot<TPR>(A, i, x, G.lane);
        if (G.lane == 0) r[i] = b[i] - ax;
    }
}

template <int TPR>
__global__ void __launch_bounds__(kBlock) vc_spmv(int64_t n, CsrOp<double> A, const double* __restrict__ x,
                                                  double* __restrict__ y, const int* done) {
    if (done && *done) return;
    RowGroups<TPR> G;
    for (int64_t i = G.first; i < n; i += G.stride) {
        const double ax = row_dot<TPR>(A, i, x, G.lane);
        if (G.lane == 0) y[i] = ax;
    }
}

template <int TPR>
__global__ void __launch_bounds__(kBlock) vc_prolong(int64_t n, CsrOp<double> P, const double* __restrict__ e,
                                                     double* __restrict__ x, const int* done) {
    if (done && *done) return;
    RowGroups<TPR> G;
    for (int64_t i = G.first; i < n; i += G.stride) {
        const double pe = row_dot<TPR>(P, i, e, G.lane);
        if (G.lane == 0) x[i] = x[i] + pe;
    }
}

__global__ void __launch_bounds__(kBlock) vc_copy(int64_t n, const double* __restrict__ a, double* __restrict__ b,
                                                  const int* done) {
    if (done && *done) return;
    int64_t lo, hi;
    block_range(n, lo, hi);
    for (int64_t i = lo + threadIdx.x; i < hi; i += kBlock) b[i] = a[i];
}

// *mean_out = sum(z) / n
__global__ void __launch_bounds__(kBlock) k_mean(int64_t n, const double* __restrict__ z, const int* done, RedBuf rb,
                                                 double* mean_out) {
    (void)mean_out;
    if (done && *done) return;
    int64_t lo, hi;
    block_range(n, lo, hi);
    double acc[1] = {0.0};
    for (int64_t i = lo + threadIdx.x; i < hi; i += kBlock) acc[0] += z[i];
    block_partials<1>(acc, rb.partials);
}

// *mean_out = (sum of the partials) / n, unless the solve is done
__global__ void __launch_bounds__(kBlock) fin_mean(const double* partials, int nparts, int64_t n, const int* done,
                                                   double* mean_out) {
    if (done && *done) return;
    double tot[1];
    if (finalize_sum<1>(partials, nparts, tot)) *mean_out = tot[0] / (double)n;
}

// x = Minv b for the dense coarse inverse (row per wave)
__global__ void __launch_bounds__(kBlock) vc_coarse(int64_t n, const double* __restrict__ M,
                                                    const double* __restrict__ b, double* __restrict__ x,
                                                    const int* done) {
    if (done && *done) return;
    const int lane = threadIdx.x % kWave;
    const int64_t row = (int64_t)blockIdx.x * (kBlock / kWave) + threadIdx.x / kWave;
    if (row >= n) return;
    double s = 0.0;
    for (int64_t j = lane; j < n; j += kWave) s += M[row * n + j] * b[j];
    s = wave_sum(s);
    if (lane == 0) x[row] = s;
}

#define TPR_LAUNCH(tpr, KERNEL, n, st, ...)                                                              \
    do {                                                                                                 \
        switch (tpr) {                                                                                   \
            case 1: KERNEL<1><<<grid_rows(n, 1), kBlock, 0, st>>>(__VA_ARGS__); break;                    \
            case 4: KERNEL<4><<<grid_rows(n, 4), kBlock, 0, st>>>(__VA_ARGS__); break;                    \
            case 16: KERNEL<16><<<grid_rows(n, 16), kBlock, 0, st>>>(__VA_ARGS__); break;                 \
            default: KERNEL<64><<<grid_rows(n, 64), kBlock, 0, st>>>(__VA_ARGS__); break;                 \
        }                                                                                                \
    } while (0)

template <int TPR>
static void launch_jacobi(int64_t n, CsrOp<double> A, const double* Dinv, const double* b, const double* x,
                          double* y, double w, const int* done, RedBuf* rb, double* mean_out, hipStream_t st) {
    const int g = grid_rows(n, TPR);
    if (rb) {
        vc_jacobi<TPR, true><<<g, kBlock, 0, st>>>(n, A, Dinv, b, x, y, w, done, *rb, mean_out);
        fin_mean<<<1, kBlock, 0, st>>>(rb->partials, g, n, done, mean_out);
    } else {
        vc_jacobi<TPR, false><<<g, kBlock, 0, st>>>(n, A, Dinv, b, x, y, w, done, RedBuf{}, nullptr);
    }
}

static void jacobi(int tpr, int64_t n, CsrOp<double> A, const double* Dinv, const double* b, const double* x,
                   double* y, double w, const int* done, RedBuf* rb, double* mean_out, hipStream_t st) {
    switch (tpr) {
        case 1: launch_jacobi<1>(n, A, Dinv, b, x, y, w, done, rb, mean_out, st); break;
        case 4: launch_jacobi<4>(n, A, Dinv, b, x, y, w, done, rb, mean_out, st); break;
        case 16: launch_jacobi<16>(n, A, Dinv, b, x, y, w, done, rb, mean_out, st); break;
        default: launch_jacobi<64>(n, A, Dinv, b, x, y, w, done, rb, mean_out, st); break;
    }
}

// Enqueue one V-cycle on level k: x_k = M_k^{-1} b_k. When `rb` is given, the sum of the
// level-0 output is reduced into *mean_out (divided by n).
static void vcycle(const Amg& M, int k, hipStream_t st, const int* done, RedBuf* rb, double* mean_out) {
    if (k == M.nlev) {
        const int rows_per_block = kBlock / kWave;
        vc_coarse<<<(int)((M.nc + rows_per_block - 1) / rows_per_block), kBlock, 0, st>>>(M.nc, M.cinv, M.cb, M.cx,
                                                                                          done);
        return;
    }
    const Level& L = M.lv[k];
    const double w = M.omega;
    double* next_b = (k + 1 == M.nlev) ? M.cb : M.lv[k + 1].b;
    double* next_x = (k + 1 == M.nlev) ? M.cx : M.lv[k + 1].x;
    const int64_t next_n = (k + 1 == M.nlev) ? M.nc : M.lv[k + 1].n;
    // pre-smoothing: `sweeps` weighted-Jacobi sweeps from zero
    double* xa = L.x;
    double* xb = L.y;
    if (M.sweeps >= 2) {
        TPR_LAUNCH(L.tpr_a, vc_pre2, L.n, st, L.n, L.A, L.Dinv, L.b, xa, w, done);
        for (int s = 2; s < M.sweeps; ++s) {
            jacobi(L.tpr_a, L.n, L.A, L.Dinv, L.b, xa, xb, w, done, nullptr, nullptr, st);
            double* t = xa;
            xa = xb;
            xb = t;
        }
    } else {
        vc_pre1<<<grid_for(L.n), kBlock, 0, st>>>(L.n, L.Dinv, L.b, xa, w, done);
    }
    TPR_LAUNCH(L.tpr_a, vc_resid, L.n, st, L.n, L.A, L.b, xa, L.r, done);
    TPR_LAUNCH(L.tpr_r, vc_spmv, next_n, st, next_n, L.R, L.r, next_b, done);
    vcycle(M, k + 1, st, done, nullptr, nullptr);
    TPR_LAUNCH(L.tpr_p, vc_prolong, L.n, st, L.n, L.P, next_x, xa, done);
    // post-smoothing: ping-pong between the level's two buffers, never in place
    double* cur = xa;
    double* other = (xa == L.x) ? L.y : L.x;
    for (int s = 0; s < M.sweeps; ++s) {
        const bool last = (s == M.sweeps - 1);
        jacobi(L.tpr_a, L.n, L.A, L.Dinv, L.b, cur, other, w, done, last ? rb : nullptr, mean_out, st);
        double* t = cur;
        cur = other;
        other = t;
    }
    if (cur != L.x) vc_copy<<<grid_for(L.n), kBlock, 0, st>>>(L.n, cur, L.x, done);
}

}  // namespace

extern "C" {

void* ftmi_amg_create(double omega, int sweeps) {
    Amg* M = new Amg();
    M->omega = omega;
    M->sweeps = sweeps;
    return M;
}

void ftmi_amg_destroy(void* h) { delete (Amg*)h; }

int ftmi_amg_add_level(void* h, int64_t n, const int32_t* a_ptr, const int32_t* a_col, const double* a_val,
                       const int32_t* p_ptr, const int32_t* p_col, const double* p_val, const int32_t* r_ptr,
                       const int32_t* r_col, const double* r_val, const double* dinv, double* vb, double* vx,
                       double* vy, double* vr) {
    Amg* M = (Amg*)h;
    if (M->nlev >= 24) {
        set_error_msg("ftmi_amg_add_level: at most 24 levels");
        return -1;
    }
    Level& L = M->lv[M->nlev++];
    L.n = n;
    L.A = CsrOp<double>{a_ptr, a_col, a_val};
    L.P = CsrOp<double>{p_ptr, p_col, p_val};
    L.R = CsrOp<double>{r_ptr, r_col, r_val};
    // row lengths from the CSR pointers (host copies of the last pointer entries)
    int32_t nnz_a = 0, nnz_p = 0, nnz_r = 0;
    int64_t nrows_r = 0;
    if (hipMemcpy(&nnz_a, a_ptr + n, sizeof(int32_t), hipMemcpyDeviceToHost) != hipSuccess ||
        hipMemcpy(&nnz_p, p_ptr + n, sizeof(int32_t), hipMemcpyDeviceToHost) != hipSuccess) {
        set_error_msg("ftmi_amg_add_level: cannot read row pointers");
        return -1;
    }
    L.tpr_a = pick_tpr(n, nnz_a);
    L.tpr_p = pick_tpr(n, nnz_p);
    L.nnz_p = nnz_p;
    L.tpr_r = 1;  // R = P^T has the next level's rows: set when that size is known
    if (M->nlev >= 2) M->lv[M->nlev - 2].tpr_r = pick_tpr(n, M->lv[M->nlev - 2].nnz_p);
    (void)nnz_r;
    (void)nrows_r;
    L.Dinv = dinv;
    L.b = vb;
    L.x = vx;
    L.y = vy;
    L.r = vr;
    return 0;
}

int ftmi_amg_set_coarse(void* h, int64_t n, const double* inv, double* vb, double* vx) {
    Amg* M = (Amg*)h;
    M->nc = n;
    if (M->nlev >= 1) M->lv[M->nlev - 1].tpr_r = pick_tpr(n, M->lv[M->nlev - 1].nnz_p);
    M->cinv = inv;
    M->cb = vb;
    M->cx = vx;
    return 0;
}

// x = M^{-1} b with one V-cycle (b, x: level-0 vectors, distinct from the level buffers).
int ftmi_amg_vcycle(void* h, const double* b, double* x, void* stream) {
    Amg* M = (Amg*)h;
    hipStream_t st = (hipStream_t)stream;
    if (M->nlev == 0) {
        const int rows_per_block = kBlock / kWave;
        vc_coarse<<<(int)((M->nc + rows_per_block - 1) / rows_per_block), kBlock, 0, st>>>(M->nc, M->cinv, b, x,
                                                                                           nullptr);
        FTMI_LAUNCH_CHECK("ftmi_amg_vcycle");
        return 0;
    }
    Level& L0 = M->lv[0];
    double *sb = L0.b, *sx = L0.x;
    L0.b = (double*)b;
    L0.x = x;
    vcycle(*M, 0, st, nullptr, nullptr, nullptr);
    L0.b = sb;
    L0.x = sx;
    FTMI_LAUNCH_CHECK("ftmi_amg_vcycle");
    return 0;
}

}  // extern "C"

// ====================================================================== PCG
namespace {

template <bool PROJ>
__global__ void __launch_bounds__(kBlock) pcg_init_a(int64_t n, const double* __restrict__ b, State* S, RedBuf rb) {
    int64_t lo, hi;
    block_range(n, lo, hi);
    double acc[2] = {0.0, 0.0};
    for (int64_t i = lo + threadIdx.x; i < hi; i += kBlock) {
        acc[0] += b[i];
        acc[1] += b[i] * b[i];
    }
    block_partials<2>(acc, rb.partials);
}

template <bool PROJ>
__global__ void __launch_bounds__(kBlock) fin_pcg_init_a(const double* partials, int nparts, int64_t n, State* S) {
    double tot[2];
    if (finalize_sum<2>(partials, nparts, tot)) {
        S->sc[S_BMEAN] = PROJ ? tot[0] / (double)n : 0.0;
        const double bn = sqrt(tot[1]);
        S->sc[S_BNORM] = bn != 0.0 ? bn : 1.0;
    }
}

template <bool PROJ>
__global__ void __launch_bounds__(kBlock) pcg_init_b(int64_t n, const double* __restrict__ b, double* __restrict__ r,
                                                     double* __restrict__ x, State* S, RedBuf rb, double tol) {
    const double mean = S->sc[S_BMEAN];
    int64_t lo, hi;
    block_range(n, lo, hi);
    double acc[1] = {0.0};
    for (int64_t i = lo + threadIdx.x; i < hi; i += kBlock) {
        const double ri = PROJ ? b[i] - mean : b[i];
        r[i] = ri;
        x[i] = 0.0;
        acc[0] += ri * ri;
    }
    block_partials<1>(acc, rb.partials);
}

__global__ void __launch_bounds__(kBlock) fin_pcg_init_b(const double* partials, int nparts, State* S, double tol) {
    double tot[1];
    if (finalize_sum<1>(partials, nparts, tot)) {
        const double rn = sqrt(tot[0]);
        S->sc[S_RNORM] = rn;
        S->sc[S_RES] = rn / S->sc[S_BNORM];
        if (rn / S->sc[S_BNORM] < tol) {
            S->st[I_DONE] = 1;
            S->st[I_CONV] = 1;
            S->st[I_ITERS] = 0;
        }
    }
}

// z -= mean(z) (projection); rz_new = r.z; first: p = z, rz = rz_new; else beta = rz_new / rz.
template <bool PROJ>
__global__ void __launch_bounds__(kBlock) pcg_z(int64_t n, const double* __restrict__ r, double* __restrict__ z,
                                                double* __restrict__ p, State* S, RedBuf rb, int first, int k) {
    if (S->st[I_DONE]) return;
    const double mean = S->sc[S_ZMEAN];
    int64_t lo, hi;
    block_range(n, lo, hi);
    double acc[1] = {0.0};
    for (int64_t i = lo + threadIdx.x; i < hi; i += kBlock) {
        double zi = z[i];
        if (PROJ) {
            zi = zi - mean;
            z[i] = zi;
        }
        if (first) p[i] = zi;
        acc[0] += r[i] * zi;
    }
    block_partials<1>(acc, rb.partials);
}

__global__ void __launch_bounds__(kBlock) fin_pcg_z(const double* partials, int nparts, State* S, int first, int k) {
    if (S->st[I_DONE]) return;
    double tot[1];
    if (finalize_sum<1>(partials, nparts, tot)) {
        double* sc = S->sc;
        if (first) {
            sc[S_RZ] = tot[0];
        } else if (sc[S_RZ] == 0.0) {
            S->st[I_DONE] = 1;
            S->st[I_CONV] = 0;
            S->st[I_ITERS] = k;
        } else {
            sc[S_BETA] = tot[0] / sc[S_RZ];
            sc[S_RZ] = tot[0];
        }
    }
}

__global__ void __launch_bounds__(kBlock) pcg_p(int64_t n, const double* __restrict__ z, double* __restrict__ p,
                                                const State* S) {
    if (S->st[I_DONE]) return;
    const double beta = S->sc[S_BETA];
    int64_t lo, hi;
    block_range(n, lo, hi);
    for (int64_t i = lo + threadIdx.x; i < hi; i += kBlock) p[i] = z[i] + beta * p[i];
}

template <class Op>
__global__ void __launch_bounds__(kBlock) pcg_ap(int64_t n, Op A, const double* __restrict__ p,
                                                 double* __restrict__ q, State* S, RedBuf rb, int k) {
    if (S->st[I_DONE]) return;
    int64_t lo, hi;
    block_range(n, lo, hi);
    double acc[1] = {0.0};
    for (int64_t i = lo + threadIdx.x; i < hi; i += kBlock) {
        const double qi = A.row(i, p);
        q[i] = qi;
        acc[0] += p[i] * qi;
    }
    block_partials<1>(acc, rb.partials);
}

__global__ void __launch_bounds__(kBlock) fin_pcg_ap(const double* partials, int nparts, State* S, int k) {
    if (S->st[I_DONE]) return;
    double tot[1];
    if (finalize_sum<1>(partials, nparts, tot)) {
        if (tot[0] <= 0.0) {
            S->st[I_DONE] = 1;
            S->st[I_CONV] = 0;
            S->st[I_ITERS] = k;
        } else {
            S->sc[S_ALPHA] = S->sc[S_RZ] / tot[0];
        }
    }
}

__global__ void __launch_bounds__(kBlock) pcg_xr(int64_t n, const double* __restrict__ p, const double* __restrict__ q,
                                                 double* __restrict__ x, double* __restrict__ r, State* S, RedBuf rb,
                                                 double tol, int k) {
    if (S->st[I_DONE]) return;
    const double a = S->sc[S_ALPHA];
    int64_t lo, hi;
    block_range(n, lo, hi);
    double acc[1] = {0.0};
    for (int64_t i = lo + threadIdx.x; i < hi; i += kBlock) {
        x[i] += a * p[i];
        const double ri = r[i] - a * q[i];
        r[i] = ri;
        acc[0] += ri * ri;
    }
    block_partials<1>(acc, rb.partials);
}

__global__ void __launch_bounds__(kBlock) fin_pcg_xr(const double* partials, int nparts, State* S, double tol, int k) {
    if (S->st[I_DONE]) return;
    double tot[1];
    if (finalize_sum<1>(partials, nparts, tot)) {
        const double rn = sqrt(tot[0]);
        S->sc[S_RNORM] = rn;
        S->sc[S_RES] = rn / S->sc[S_BNORM];
        if (rn / S->sc[S_BNORM] < tol) {
            S->st[I_DONE] = 1;
            S->st[I_CONV] = 1;
            S->st[I_ITERS] = k;
        }
    }
}

struct PcgWork {
    double *r, *z, *p, *q;
    State* S;
    RedBuf rb;
};

static PcgWork carve_pcg(int64_t n, void* work) {
    char* w = (char*)work;
    PcgWork W;
    double* d = (double*)w;
    W.r = d;
    W.z = d + n;
    W.p = d + 2 * n;
    W.q = d + 3 * n;
    char* tail = w + ((4 * n * sizeof(double) + 255) / 256) * 256;
    W.S = (State*)tail;
    W.rb.counter = (unsigned*)(tail + 256);
    W.rb.partials = (double*)(tail + 512);
    return W;
}

template <class Op>
int pcg_run(int64_t n, Op A, const double* b, double* x, double tol, int64_t maxiter, void* amg, int project,
            void* work, int64_t* iters, double* res, hipStream_t st) {
    Pinned* pn = pinned();
    if (!pn) {
        set_error_msg("pcg: pinned host memory unavailable");
        return -1;
    }
    Amg* M = (Amg*)amg;
    PcgWork W = carve_pcg(n, work);
    const int g = grid_for(n);
    const int* done = &W.S->st[I_DONE];
    k_state_reset<<<1, 64, 0, st>>>(W.S, W.rb.counter);
    if (project) {
        pcg_init_a<true><<<g, kBlock, 0, st>>>(n, b, W.S, W.rb);
        fin_pcg_init_a<true><<<1, kBlock, 0, st>>>(W.rb.partials, g, n, W.S);
        pcg_init_b<true><<<g, kBlock, 0, st>>>(n, b, W.r, x, W.S, W.rb, tol);
    } else {
        pcg_init_a<false><<<g, kBlock, 0, st>>>(n, b, W.S, W.rb);
        fin_pcg_init_a<false><<<1, kBlock, 0, st>>>(W.rb.partials, g, n, W.S);
        pcg_init_b<false><<<g, kBlock, 0, st>>>(n, b, W.r, x, W.S, W.rb, tol);
    }
    fin_pcg_init_b<<<1, kBlock, 0, st>>>(W.rb.partials, g, W.S, tol);
    auto precond = [&]() {
        bool fused_mean = false;
        if (M && M->nlev > 0) {
            Level& L0 = M->lv[0];
            double *sb = L0.b, *sx = L0.x;
            L0.b = W.r;
            L0.x = W.z;
            vcycle(*M, 0, st, done, project ? &W.rb : nullptr, &W.S->sc[S_ZMEAN]);
            L0.b = sb;
            L0.x = sx;
            fused_mean = project != 0;
        } else if (M) {
            const int rpb = kBlock / kWave;
            vc_coarse<<<(int)((M->nc + rpb - 1) / rpb), kBlock, 0, st>>>(M->nc, M->cinv, W.r, W.z, done);
        } else {
            vc_copy<<<g, kBlock, 0, st>>>(n, W.r, W.z, done);
        }
        if (project && !fused_mean) {
            k_mean<<<g, kBlock, 0, st>>>(n, W.z, done, W.rb, &W.S->sc[S_ZMEAN]);
            fin_mean<<<1, kBlock, 0, st>>>(W.rb.partials, g, n, done, &W.S->sc[S_ZMEAN]);
        }
    };
    precond();
    if (project)
        pcg_z<true><<<g, kBlock, 0, st>>>(n, W.r, W.z, W.p, W.S, W.rb, 1, 0);
    else
        pcg_z<false><<<g, kBlock, 0, st>>>(n, W.r, W.z, W.p, W.S, W.rb, 1, 0);
    fin_pcg_z<<<1, kBlock, 0, st>>>(W.rb.partials, g, W.S, 1, 0);
    FTMI_LAUNCH_CHECK("pcg init");
    if (post_state(pn, 0, W.S, st)) return -1;
    int last = 0;
    for (int64_t k = 1; k <= maxiter; ++k) {
        pcg_ap<Op><<<g, kBlock, 0, st>>>(n, A, W.p, W.q, W.S, W.rb, (int)k);
        fin_pcg_ap<<<1, kBlock, 0, st>>>(W.rb.partials, g, W.S, (int)k);
        pcg_xr<<<g, kBlock, 0, st>>>(n, W.p, W.q, x, W.r, W.S, W.rb, tol, (int)k);
        fin_pcg_xr<<<1, kBlock, 0, st>>>(W.rb.partials, g, W.S, tol, (int)k);
        precond();
        if (project)
            pcg_z<true><<<g, kBlock, 0, st>>>(n, W.r, W.z, W.p, W.S, W.rb, 0, (int)k);
        else
            pcg_z<false><<<g, kBlock, 0, st>>>(n, W.r, W.z, W.p, W.S, W.rb, 0, (int)k);
        fin_pcg_z<<<1, kBlock, 0, st>>>(W.rb.partials, g, W.S, 0, (int)k);
        pcg_p<<<g, kBlock, 0, st>>>(n, W.z, W.p, W.S);
        FTMI_LAUNCH_CHECK("pcg iteration");
        if (post_state(pn, last ^ 1, W.S, st)) return -1;
        if (wait_state(pn, last)) return -1;
        const bool prev_done = pn->h[last].st[I_DONE] != 0;
        last ^= 1;
        if (prev_done) break;
    }
    if (wait_state(pn, last)) return -1;
    const int slot = last;
    const State& H = pn->h[slot];
    if (H.st[I_DONE]) {
        *iters = H.st[I_ITERS];
        *res = H.sc[S_RES];
        return H.st[I_CONV];
    }
    *iters = maxiter;
    *res = H.sc[S_RES];
    return 0;
}

}  // namespace

extern "C" {

int64_t ftmi_pcg_work_bytes(int64_t n) {
    return ((4 * n * (int64_t)sizeof(double) + 255) / 256) * 256 + 512 + (int64_t)kMaxGrid * kMaxRed * 8;
}

int ftmi_pcg_ldu(int64_t n, const double* diag, const int32_t* ptr, const int32_t* nbr, const double* off,
                 const double* b, double* x, double tol, int64_t maxiter, void* amg, int project, void* work,
                 int64_t* iters, double* res, void* stream) {
    LduOp<double> A{diag, ptr, nbr, off};
    return pcg_run(n, A, b, x, tol, maxiter, amg, project, work, iters, res, (hipStream_t)stream);
}

int ftmi_pcg_csr(int64_t n, const int32_t* indptr, const int32_t* indices, const double* data, const double* b,
                 double* x, double tol, int64_t maxiter, void* amg, int project, void* work, int64_t* iters,
                 double* res, void* stream) {
    CsrOp<double> A{indptr, indices, data};
    return pcg_run(n, A, b, x, tol, maxiter, amg, project, work, iters, res, (hipStream_t)stream);
}

}  // extern "C"
