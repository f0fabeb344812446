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
// The host checks `done` one iteration behind: the finalize kernel of each iteration writes the
// state block into coherent pinned host memory and stamps it with a sequence number, which the
// host polls, so the GPU never waits for the host and no copy or event is queued per iteration.
//
// The BiCGSTAB iterate lives in two buffers (the caller's x and a work buffer): the update
// x + alpha phat + omega shat is written to the buffer that is not current and becomes
// current only when ||r|| is finite, which is the reference's "keep the last finite
// iterate" rule without an extra pass.
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <vector>

#include <hip/hip_ext.h>

#include "../../include/ftmi.h"
#include "ftmi_common.h"
#include "ftmi_ops.h"

using namespace ftmi;

namespace {

// ------------------------------------------------------------------ device state block
// S_RZ, S_RZ + 1: r.z of odd / even PCG iterations (the kernel that computes the new value
// reads the previous one, so they live in different slots).
// Rule for every kernel below: workgroup 0 may write a state slot only if no workgroup of the
// same kernel reads that slot (workgroups that start later would see the new value).
enum { S_RHO, S_ALPHA, S_OMEGA, S_RNORM, S_BNORM, S_BABS, S_RHONEW, S_RES, S_RZ, S_RZ1, S_ZMEAN, S_BMEAN, S_TOLIN,
       S_NS = 16 };
enum { I_DONE, I_CONV, I_ITERS, I_CUR, I_FIXUP, I_NI = 8 };

struct State {  // 192 bytes; `seq` stamps a host copy (see post_host)
    double sc[S_NS];
    int st[I_NI];
    unsigned seq;
    int pad[7];
};

// Report of a capped solve: ok if it finished (done) or ran its full maxiter (full).
__global__ void k_report(const State* S, double* out, int full, int64_t cap) {
    if (threadIdx.x != 0) return;
    const int done = S->st[I_DONE];
    out[0] = (done || full) ? 1.0 : 0.0;
    out[1] = done ? (double)S->st[I_ITERS] : (double)cap;
    out[2] = done ? (double)S->st[I_CONV] : 0.0;
    out[3] = S->sc[S_RES];
}

__global__ void k_state_reset(State* s, unsigned* counter) {
    const int t = threadIdx.x;
    if (t < S_NS) s->sc[t] = 0.0;
    if (t < I_NI) s->st[t] = 0;
    if (t == 0) *counter = 0u;
}

// ------------------------------------------------------------------ host-side flag reader
// Two host slots in coherent (fine-grained) pinned memory: finalize kernels store the state
// block there directly and stamp it, the host polls the stamp.
struct Pinned {
    State* h = nullptr;  // host view
    State* d = nullptr;  // device view of the same memory
    unsigned want[2] = {0u, 0u};
    unsigned seq = 0u;
    bool ok = false;
    Pinned() {
        if (hipHostMalloc((void**)&h, 2 * sizeof(State), hipHostMallocCoherent | hipHostMallocMapped) != hipSuccess)
            return;
        std::memset(h, 0, 2 * sizeof(State));
        if (hipHostGetDevicePointer((void**)&d, h, 0) != hipSuccess) return;
        ok = true;
    }
};
static thread_local Pinned* g_pinned = nullptr;
static Pinned* pinned() {
    if (!g_pinned) g_pinned = new Pinned();
    return g_pinned->ok ? g_pinned : nullptr;
}

// Where the next finalize kernel posts: the device address of `slot` and a fresh stamp.
struct Post {
    State* m;
    unsigned seq;
};
static Post next_post(Pinned* pn, int slot) {
    pn->want[slot] = ++pn->seq;
    if (pn->seq == 0u) pn->want[slot] = ++pn->seq;  // 0 is the cleared value
    return Post{pn->d + slot, pn->want[slot]};
}
static constexpr Post kNoPost{nullptr, 0u};

// Thread 0 of a finalize kernel: copy the state block to the host slot, then (after a
// system-scope fence) its stamp, so a host that sees the stamp sees the fields.
__device__ inline void post_host(const State* S, Post P) {
    if (!P.m || threadIdx.x != 0) return;
    double sc[S_NS];  // all loads first, then all stores: one wait instead of one per field
    int st[I_NI];
#pragma unroll
    for (int j = 0; j < S_NS; ++j) sc[j] = S->sc[j];
#pragma unroll
    for (int j = 0; j < I_NI; ++j) st[j] = S->st[j];
#pragma unroll
    for (int j = 0; j < S_NS; ++j) P.m->sc[j] = sc[j];
#pragma unroll
    for (int j = 0; j < I_NI; ++j) P.m->st[j] = st[j];
    __threadfence_system();
    __hip_atomic_store(&P.m->seq, P.seq, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
}

// Spin until slot `slot` carries its stamp. The stream is queried now and then, so a failed
// launch is reported and a drained stream without the stamp is an error instead of a hang.
static int wait_state(Pinned* pn, int slot, hipStream_t st) {
    const unsigned want = pn->want[slot];
    const volatile unsigned* p = &pn->h[slot].seq;
    for (unsigned spin = 1u;; ++spin) {
        if (*p == want) break;
        if ((spin & 255u) == 0u) {
            const hipError_t e = hipStreamQuery(st);
            if (e == hipSuccess) {
                if (*p == want) break;
                set_error_msg("krylov: the solver state was not posted");
                return -1;
            }
            if (e != hipErrorNotReady) {
                set_error_msg("krylov: stream error while waiting for the solver state");
                return -1;
            }
        }
        __builtin_ia32_pause();
    }
    __atomic_thread_fence(__ATOMIC_ACQUIRE);
    return 0;
}

// ------------------------------------------------------------------ opt-in kernel timer
// When enabled, the launches of one kernel (the BiCGSTAB p-update + SpMV, the solver's
// largest share of time) carry a start and a stop event in their dispatch
// (hipExtLaunchKernelGGL), so a benchmark can read that kernel's average duration inside a
// real solve, as a kernel trace would.
struct KTimer {
    bool on = false;
    std::vector<std::pair<hipEvent_t, hipEvent_t>> pending;
    std::vector<hipEvent_t> pool;
    double total_ms = 0.0;
    int64_t launches = 0;
    hipEvent_t get() {
        if (!pool.empty()) {
            hipEvent_t e = pool.back();
            pool.pop_back();
            return e;
        }
        // device-scope release: the event does not write back and invalidate caches
        // system-wide, so the bracket measures the kernel rather than a cache flush
        hipEvent_t e = nullptr;
        hipEventCreateWithFlags(&e, hipEventReleaseToDevice);
        return e;
    }
    void drain() {
        for (auto& pr : pending) {
            float ms = 0.f;
            if (hipEventSynchronize(pr.second) == hipSuccess && hipEventElapsedTime(&ms, pr.first, pr.second) == hipSuccess) {
                total_ms += ms;
                ++launches;
            }
            pool.push_back(pr.first);
            pool.push_back(pr.second);
        }
        pending.clear();
    }
};
static KTimer g_timer;

// Capture mode: the solvers never read the device state back. Every iteration up to its cap
// is enqueued (kernels of a finished solve return at once on the device `done` flag) and the
// final fix-ups read the state on the device, so a solve can be recorded in a HIP graph.
// The iteration counts and residuals are then not reported (*iters = -1, *res = NaN).
static bool g_capture = false;

// Capture plan (ftmi_capture_plan): the solves recorded in capture mode take slots 0, 1, ...
// in call order; slot s enqueues min(maxiter, caps[s]) iterations and, when it ends, writes
// its report (4 doubles: ok, iterations, converged, residual) to report + 4 s on the device.
// ok = 1 when the recorded iterations gave the eager solve's result (the solve finished
// within its cap, or the cap was maxiter); the host redoes the step eagerly otherwise.
struct CapturePlan {
    std::vector<int64_t> caps;
    double* report = nullptr;
    int slots = 0;
    int next = 0;
};
static CapturePlan g_plan;

// The next slot's cap and report address (nullptr: no plan, or no room left).
static double* capture_slot(int64_t maxiter, int64_t* cap) {
    *cap = maxiter;
    if (!g_capture || !g_plan.report) return nullptr;
    const int s = g_plan.next++;
    if (s < (int)g_plan.caps.size() && g_plan.caps[s] > 0 && g_plan.caps[s] < maxiter) *cap = g_plan.caps[s];
    return s < g_plan.slots ? g_plan.report + 4 * s : nullptr;
}

// ------------------------------------------------------------------ BiCGSTAB kernels
// One iteration is three launches: the p-update fused into v = A phat (phat recomputed at the
// neighbours from r, p, v of the previous iteration), the s-update fused into t = A shat, and
// the residual/iterate update. The scalars of a reduction are computed by every workgroup of
// the kernel that consumes it (consume_sum); the end-of-iteration bookkeeping (norms, flags,
// the host post) is done by the next iteration's first kernel, whose workgroups all take the
// same decisions from the same partials and of which workgroup 0 writes the state. phat and
// shat are recomputed as inv p and inv s where they are needed again (bit-identical to
// storing them), so the iteration moves two vectors fewer.
// ZERO: the start vector is zero (x is cleared and r = b without the product)
template <class Op, typename T, bool ZERO>
__global__ void __launch_bounds__(kBlock) bi_init(int64_t n, Op A, const T* __restrict__ b, T* __restrict__ x,
                                                  T* __restrict__ r, T* __restrict__ rh, T* __restrict__ p,
                                                  T* __restrict__ v, double* partials) {
    int64_t lo, hi;
    block_range(n, lo, hi);
    T acc[2] = {T(0), T(0)};
    for (int64_t i = lo + threadIdx.x; i < hi; i += kBlock) {
        T ri;
        if (ZERO) {
            x[i] = T(0);
            ri = b[i];
        } else {
            ri = b[i] - A.row(i, x);
        }
        r[i] = ri;
        rh[i] = ri;
        p[i] = T(0);
        v[i] = T(0);
        acc[0] += ri * ri;
        acc[1] += b[i] * b[i];
    }
    block_partials<2>(acc, partials);
}

__global__ void __launch_bounds__(kBlock) fin_bi_init(const double* partials, int nparts, State* S, double tol,
                                                      int rel, Post post) {
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
    post_host(S, post);
}

// End of iteration k from the totals (r.r, rhat.r) of its residual update: the iterate
// becomes current when ||r|| is finite (else the solve stops on the last finite one), then
// the convergence test, omega = 0 and rho = 0 breakdowns. Every thread that calls it gets the
// same answer (true: the solve is done); only `lead` writes the state block.
// State words it reads that no kernel of the iteration in flight writes (loaded by the caller,
// ahead of the partial sums)
struct BiScal {
    double bnorm, babs, omega;
};

__device__ inline bool bi_finish_iteration(State* S, const double (&tot)[2], double tol, int k, int maxiter,
                                           bool lead, const BiScal& v) {
    double* sc = S->sc;
    int* st = S->st;
    const double rnorm = sqrt(tot[0]);
    if (!isfinite(rnorm)) {  // breakdown: the current iterate stays
        if (lead) {
            st[I_DONE] = 1;
            st[I_CONV] = 0;
            st[I_ITERS] = k;
            sc[S_RES] = __builtin_nan("");
        }
        return true;
    }
    const double bnorm = v.bnorm;
    if (lead) {
        st[I_CUR] ^= 1;
        sc[S_RNORM] = rnorm;
        sc[S_RES] = rnorm / bnorm;
    }
    int done = 0, conv = 0, iters = k;
    if (rnorm / bnorm < tol || rnorm <= 1e-14 * v.babs) {
        done = conv = 1;
    } else if (v.omega == 0.0) {
        done = 1;
    } else {
        if (lead) sc[S_RHONEW] = tot[1];
        if (tot[1] == 0.0 && k < maxiter) {  // the next iteration breaks at its first step
            done = 1;
            iters = k + 1;
        }
    }
    if (done && lead) {
        st[I_DONE] = 1;
        st[I_CONV] = conv;
        st[I_ITERS] = iters;
    }
    return done != 0;
}

// The end of iteration k as a kernel of its own (after the last enqueued iteration, and in
// capture mode's tail).
__global__ void __launch_bounds__(kBlock) fin_bi_r(const double* partials, int nparts, State* S, double tol, int k,
                                                   int maxiter, Post post) {
    double tot[2];
    if (!S->st[I_DONE] && finalize_sum<2>(partials, nparts, tot))
        bi_finish_iteration(S, tot, tol, k, maxiter, true, BiScal{S->sc[S_BNORM], S->sc[S_BABS], S->sc[S_OMEGA]});
    post_host(S, post);
}

// FIN: first finish iteration k - 1 from the partials of its residual update (the work of
// fin_bi_r, posting the state to the host), then, unless done:
// p = r + beta (p_prev - omega v_prev), v = A (inv p); partials of rhat . v
template <class Op, typename T, bool FIN>
__global__ void __launch_bounds__(kBlock) bi_pv(int64_t n, Op A, const T* __restrict__ r,
                                                const T* __restrict__ p_prev, const T* __restrict__ v_prev,
                                                const T* __restrict__ inv, const T* __restrict__ rh,
                                                T* __restrict__ p, T* __restrict__ v, State* S,
                                                const double* part_in, int np_in, double tol, int k, int maxiter,
                                                Post post, double* part_out) {
    const bool lead = blockIdx.x == 0 && threadIdx.x == 0;
    // every state word read here is loaded with the done flag, ahead of the partial sums, so
    // their latencies overlap (the finish of the previous iteration writes none of them)
    const double* sc = S->sc;
    const int done0 = S->st[I_DONE];
    double rhonew = sc[S_RHONEW];
    const double rho = sc[S_RHO], alpha = sc[S_ALPHA];
    const BiScal fs{sc[S_BNORM], sc[S_BABS], sc[S_OMEGA]};
    if (done0) {
        if (FIN && lead) post_host(S, post);
        return;
    }
    if (FIN) {
        double tot[2];
        consume_sum<2>(part_in, np_in, tot);
        const bool done = bi_finish_iteration(S, tot, tol, k - 1, maxiter, lead, fs);
        if (lead) post_host(S, post);
        if (done) return;
        rhonew = tot[1];
    }
    const T beta = (T)((rhonew / rho) * (alpha / fs.omega));
    const T om = (T)fs.omega;
    auto phat = [&](int64_t j) { return inv[j] * (r[j] + beta * (p_prev[j] - om * v_prev[j])); };
    int64_t lo, hi;
    block_range(n, lo, hi);
    T acc[1] = {T(0)};
    for (int64_t i = lo + threadIdx.x; i < hi; i += kBlock) {
        const T pi = r[i] + beta * (p_prev[i] - om * v_prev[i]);
        p[i] = pi;
        const T vi = A.row_f(i, phat);
        v[i] = vi;
        acc[0] += rh[i] * vi;
    }
    block_partials<1>(acc, part_out);
}

// alpha = rho / (rhat . v); s = r - alpha v, t = A (inv s); partials s.s, t.t, t.s
template <class Op, typename T>
__global__ void __launch_bounds__(kBlock) bi_st(int64_t n, Op A, const T* __restrict__ r, const T* __restrict__ v,
                                                const T* __restrict__ inv, T* __restrict__ s, T* __restrict__ t,
                                                State* S, const double* part_in, int np_in, double* part_out, int k) {
    // the state words this kernel reads are loaded together with the done flag, ahead of the
    // partial sums, so their latencies overlap
    const int done = S->st[I_DONE];
    const double rho = S->sc[S_RHONEW];
    if (done) return;
    double tot[1];
    consume_sum<1>(part_in, np_in, tot);
    const bool lead = blockIdx.x == 0 && threadIdx.x == 0;
    if (tot[0] == 0.0) {  // breakdown: the iteration stops before its first update
        if (lead) {
            S->sc[S_RHO] = rho;
            S->st[I_DONE] = 1;
            S->st[I_CONV] = 0;
            S->st[I_ITERS] = k;
            S->sc[S_RES] = S->sc[S_RNORM] / S->sc[S_BNORM];
        }
        return;
    }
    const double alpha_d = rho / tot[0];
    if (lead) {
        S->sc[S_RHO] = rho;
        S->sc[S_ALPHA] = alpha_d;
    }
    const T alpha = (T)alpha_d;
    auto shat = [&](int64_t j) { return inv[j] * (r[j] - alpha * v[j]); };
    int64_t lo, hi;
    block_range(n, lo, hi);
    T acc[3] = {T(0), T(0), T(0)};
    for (int64_t i = lo + threadIdx.x; i < hi; i += kBlock) {
        const T si = r[i] - alpha * v[i];
        s[i] = si;
        const T ti = A.row_f(i, shat);
        t[i] = ti;
        acc[0] += si * si;
        acc[1] += ti * ti;
        acc[2] += ti * si;
    }
    block_partials<3>(acc, part_out);
}

// early exit on ||s||; omega = (t.s)/(t.t); r = s - omega t;
// x_other = x_cur + alpha (inv p) + omega (inv s)
template <typename T>
__global__ void __launch_bounds__(kBlock) bi_r(int64_t n, const T* __restrict__ s, const T* __restrict__ t,
                                               const T* __restrict__ p, const T* __restrict__ inv,
                                               const T* __restrict__ rh, T* __restrict__ r, T* __restrict__ x0,
                                               T* __restrict__ x1, State* S, const double* part_in, int np_in,
                                               double* part_out, double tol, int k) {
    const int done = S->st[I_DONE];
    const double bnorm = S->sc[S_BNORM], alpha_d = S->sc[S_ALPHA];
    const int cur = S->st[I_CUR];
    if (done) return;
    double tot[3];
    consume_sum<3>(part_in, np_in, tot);
    const bool lead = blockIdx.x == 0 && threadIdx.x == 0;
    const double snorm = sqrt(tot[0]);
    if (snorm / bnorm < tol) {  // x = x_cur + alpha phat, applied after the loop
        if (lead) {
            S->st[I_DONE] = 1;
            S->st[I_CONV] = 1;
            S->st[I_ITERS] = k;
            S->st[I_FIXUP] = 1;
            S->sc[S_RES] = snorm / bnorm;
        }
        return;
    }
    const double om_d = tot[1] > 0.0 ? tot[2] / tot[1] : 0.0;
    if (lead) S->sc[S_OMEGA] = om_d;
    const T om = (T)om_d;
    const T alpha = (T)alpha_d;
    const T* xc = cur ? x1 : x0;
    T* xo = cur ? x0 : x1;
    int64_t lo, hi;
    block_range(n, lo, hi);
    T acc[2] = {T(0), T(0)};
    for (int64_t i = lo + threadIdx.x; i < hi; i += kBlock) {
        const T si = s[i];
        const T ri = si - om * t[i];
        r[i] = ri;
        const T wi = inv[i];
        xo[i] = xc[i] + alpha * (wi * p[i]) + om * (wi * si);
        acc[0] += ri * ri;
        acc[1] += rh[i] * ri;
    }
    block_partials<2>(acc, part_out);
}

// x0 = x_cur + alpha (inv p) (early exit on ||s||), or x0 = x1 when x1 is current.
template <typename T>
__global__ void __launch_bounds__(kBlock) bi_final(int64_t n, T* __restrict__ x0, const T* __restrict__ x1,
                                                   const T* __restrict__ p, const T* __restrict__ inv,
                                                   double alpha_d, int cur, int fixup) {
    const T alpha = (T)alpha_d;
    int64_t lo, hi;
    block_range(n, lo, hi);
    for (int64_t i = lo + threadIdx.x; i < hi; i += kBlock) {
        const T xc = cur ? x1[i] : x0[i];
        x0[i] = fixup ? xc + alpha * (inv[i] * p[i]) : xc;
    }
}

// bi_final with cur / fixup / alpha read from the state block on the device; p of the last
// iteration is p_odd or p_even by the parity of the iteration count
template <typename T>
__global__ void __launch_bounds__(kBlock) bi_final_dev(int64_t n, T* __restrict__ x0, const T* __restrict__ x1,
                                                       const T* __restrict__ p_even, const T* __restrict__ p_odd,
                                                       const T* __restrict__ inv, const State* S) {
    const int cur = S->st[I_CUR], fixup = S->st[I_FIXUP];
    if (!cur && !fixup) return;
    const T alpha = (T)S->sc[S_ALPHA];
    const T* p = (S->st[I_ITERS] & 1) ? p_odd : p_even;
    int64_t lo, hi;
    block_range(n, lo, hi);
    for (int64_t i = lo + threadIdx.x; i < hi; i += kBlock) {
        const T xc = cur ? x1[i] : x0[i];
        x0[i] = fixup ? xc + alpha * (inv[i] * p[i]) : xc;
    }
}

template <typename T>
struct BiWork {
    T *r, *rh, *p[2], *v[2], *s, *t, *x1;
    State* S;
    unsigned* counter;
    double* part[3];
};

constexpr int64_t kPartBytes = (int64_t)kRedGrid * kMaxRed * 8;
inline int64_t align256(int64_t b) { return (b + 255) / 256 * 256; }
constexpr int kBiVectors = 9;

template <typename T>
static int64_t bi_bytes(int64_t n) {
    return align256(kBiVectors * n * (int64_t)sizeof(T)) + 512 + 3 * kPartBytes;
}

template <typename T>
static BiWork<T> carve_bi(int64_t n, void* work) {
    char* w = (char*)work;
    BiWork<T> W;
    T* d = (T*)w;
    W.r = d;
    W.rh = d + n;
    W.p[0] = d + 2 * n;
    W.p[1] = d + 3 * n;
    W.v[0] = d + 4 * n;
    W.v[1] = d + 5 * n;
    W.s = d + 6 * n;
    W.t = d + 7 * n;
    W.x1 = d + 8 * n;
    char* tail = w + align256(kBiVectors * n * (int64_t)sizeof(T));
    W.S = (State*)tail;
    W.counter = (unsigned*)(tail + 256);
    for (int j = 0; j < 3; ++j) W.part[j] = (double*)(tail + 512 + j * kPartBytes);
    return W;
}

// the mixed-precision transport solve keeps its FP32 right-hand side and correction in
// front of an FP32 BiCGSTAB workspace
static int64_t bi_mixed_bytes(int64_t n) { return align256(2 * n * (int64_t)sizeof(float)) + bi_bytes<float>(n); }

// Enqueue iteration k: bi_pv (finishing iteration k - 1 when k > 1, posting its state to
// `post`), bi_st, bi_r. Partials: bi_r -> part[2] -> bi_pv -> part[0] -> bi_st -> part[1]
// -> bi_r, so no kernel reads a buffer it writes.
template <class Op, typename T>
static void bi_iteration(int64_t n, Op A, const T* inv, T* x, BiWork<T>& W, int g, double tol, int64_t k,
                         int64_t maxiter, Post post, hipStream_t st, hipEvent_t t0 = nullptr,
                         hipEvent_t t1 = nullptr) {
    const int a = (int)((k - 1) & 1), c = (int)(k & 1);
    const T* r = W.r;
    if (k > 1) {
        if (t0)  // events carried by the dispatch: the kernel's own start and end
            hipExtLaunchKernelGGL(bi_pv<Op, T, true>, dim3(g), dim3(kBlock), 0, st, t0, t1, 0, n, A, r,
                                  (const T*)W.p[a], (const T*)W.v[a], inv, (const T*)W.rh, W.p[c], W.v[c], W.S,
                                  (const double*)W.part[2], g, tol, (int)k, (int)maxiter, post, W.part[0]);
        else
            bi_pv<Op, T, true><<<g, kBlock, 0, st>>>(n, A, r, W.p[a], W.v[a], inv, W.rh, W.p[c], W.v[c], W.S,
                                                     W.part[2], g, tol, (int)k, (int)maxiter, post, W.part[0]);
    } else {
        if (t0)
            hipExtLaunchKernelGGL(bi_pv<Op, T, false>, dim3(g), dim3(kBlock), 0, st, t0, t1, 0, n, A, r,
                                  (const T*)W.p[a], (const T*)W.v[a], inv, (const T*)W.rh, W.p[c], W.v[c], W.S,
                                  (const double*)nullptr, 0, tol, (int)k, (int)maxiter, kNoPost, W.part[0]);
        else
            bi_pv<Op, T, false><<<g, kBlock, 0, st>>>(n, A, r, W.p[a], W.v[a], inv, W.rh, W.p[c], W.v[c], W.S,
                                                      nullptr, 0, tol, (int)k, (int)maxiter, kNoPost, W.part[0]);
    }
    bi_st<Op, T><<<g, kBlock, 0, st>>>(n, A, W.r, W.v[c], inv, W.s, W.t, W.S, W.part[0], g, W.part[1], (int)k);
    bi_r<T><<<g, kBlock, 0, st>>>(n, W.s, W.t, W.p[c], inv, W.rh, W.r, x, W.x1, W.S, W.part[1], g, W.part[2], tol,
                                  (int)k);
}

}  // namespace

extern "C" int64_t ftmi_bicgstab_work_bytes(int64_t n) {
    const int64_t a = bi_bytes<double>(n), b = bi_mixed_bytes(n);
    return a > b ? a : b;
}

namespace ftmi {

// Runs the whole solve; returns 1 if converged, 0 if not, <0 on a HIP error.
// zero_x0: x is cleared and taken as the start vector (no product in the initial residual)
template <class Op, typename T>
int bicgstab_run(int64_t n, Op A, const T* b, const T* inv, T* x, double tol, int64_t maxiter, int rel, void* work,
                 int64_t* iters, double* res, hipStream_t st, bool zero_x0) {
    Pinned* pn = pinned();
    if (!pn) {
        set_error_msg("bicgstab: pinned host memory unavailable");
        return -1;
    }
    BiWork<T> W = carve_bi<T>(n, work);
    const int g = grid_red(n);
    k_state_reset<<<1, 64, 0, st>>>(W.S, W.counter);
    if (zero_x0)
        bi_init<Op, T, true><<<g, kBlock, 0, st>>>(n, A, b, x, W.r, W.rh, W.p[0], W.v[0], W.part[0]);
    else
        bi_init<Op, T, false><<<g, kBlock, 0, st>>>(n, A, b, x, W.r, W.rh, W.p[0], W.v[0], W.part[0]);
    fin_bi_init<<<1, kBlock, 0, st>>>(W.part[0], g, W.S, tol, rel, g_capture ? kNoPost : next_post(pn, 0));
    FTMI_LAUNCH_CHECK("bicgstab init");
    if (g_capture) {
        int64_t cap;
        double* report = capture_slot(maxiter, &cap);
        for (int64_t k = 1; k <= cap; ++k) bi_iteration<Op, T>(n, A, inv, x, W, g, tol, k, maxiter, kNoPost, st);
        if (cap >= 1) fin_bi_r<<<1, kBlock, 0, st>>>(W.part[2], g, W.S, tol, (int)cap, (int)maxiter, kNoPost);
        bi_final_dev<T><<<grid_for(n), kBlock, 0, st>>>(n, x, W.x1, W.p[0], W.p[1], inv, W.S);
        if (report) k_report<<<1, 64, 0, st>>>(W.S, report, cap >= maxiter ? 1 : 0, cap);
        FTMI_LAUNCH_CHECK("bicgstab (capture)");
        *iters = -1;
        *res = __builtin_nan("");
        return 0;
    }
    // The state after iteration j goes to host slot j & 1 (j = 0: fin_bi_init). Iteration k's
    // first kernel posts the state after k - 1; the host then waits for the one after k - 2,
    // so the GPU always has the rest of an iteration queued while the host reads.
    int64_t k = 1;
    for (; k <= maxiter; ++k) {
        const Post post = k > 1 ? next_post(pn, (int)((k - 1) & 1)) : kNoPost;
        hipEvent_t t0 = nullptr, t1 = nullptr;
        if (g_timer.on) {
            t0 = g_timer.get();
            t1 = g_timer.get();
            g_timer.pending.emplace_back(t0, t1);
        }
        bi_iteration<Op, T>(n, A, inv, x, W, g, tol, k, maxiter, post, st, t0, t1);
        FTMI_LAUNCH_CHECK("bicgstab iteration");
        if (k >= 2) {
            const int slot = (int)(k & 1);  // the state after iteration k - 2
            if (wait_state(pn, slot, st)) return -1;
            if (pn->h[slot].st[I_DONE]) {
                ++k;
                break;
            }
        }
    }
    const int64_t last = k - 1;  // the last iteration enqueued (0: none, the state is fin_bi_init's)
    const int slot = (int)(last & 1);
    if (last > 0) {
        fin_bi_r<<<1, kBlock, 0, st>>>(W.part[2], g, W.S, tol, (int)last, (int)maxiter, next_post(pn, slot));
        FTMI_LAUNCH_CHECK("bicgstab finish");
    }
    if (wait_state(pn, slot, st)) return -1;
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
        bi_final<T><<<grid_for(n), kBlock, 0, st>>>(n, x, W.x1, W.p[H.st[I_ITERS] & 1], inv, H.sc[S_ALPHA],
                                                    H.st[I_CUR], H.st[I_FIXUP]);
    FTMI_LAUNCH_CHECK("bicgstab final");
    return conv;
}

template int bicgstab_run<CsrOp<double>, double>(int64_t, CsrOp<double>, const double*, const double*, double*,
                                                 double, int64_t, int, void*, int64_t*, double*, hipStream_t, bool);
template int bicgstab_run<LduOp<double>, double>(int64_t, LduOp<double>, const double*, const double*, double*,
                                                 double, int64_t, int, void*, int64_t*, double*, hipStream_t, bool);
template int bicgstab_run<LduSellOp<double>, double>(int64_t, LduSellOp<double>, const double*, const double*, double*,
                                                     double, int64_t, int, void*, int64_t*, double*, hipStream_t,
                                                     bool);

}  // namespace ftmi

extern "C" {

int ftmi_timer_enable(int on) {
    g_timer.drain();
    g_timer.on = on != 0;
    g_timer.total_ms = 0.0;
    g_timer.launches = 0;
    return 0;
}

int ftmi_capture_mode(int on) {
    const int was = g_capture ? 1 : 0;
    g_capture = on != 0;
    return was;
}

int ftmi_capture_plan(const int64_t* caps, int ncaps, double* report, int slots) {
    g_plan.caps.assign(caps && ncaps > 0 ? caps : nullptr, caps && ncaps > 0 ? caps + ncaps : nullptr);
    g_plan.report = report;
    g_plan.slots = report ? slots : 0;
    g_plan.next = 0;
    return 0;
}

int ftmi_capture_slots_used(void) { return g_plan.next; }

int ftmi_timer_read(double* total_ms, int64_t* launches) {
    g_timer.drain();
    *total_ms = g_timer.total_ms;
    *launches = g_timer.launches;
    return 0;
}

int ftmi_bicgstab_jacobi(int64_t n, const int32_t* indptr, const int32_t* indices, const double* data,
                         const double* b, const double* inv_diag, double* x, double tol, int64_t maxiter,
                         int relative_to_initial, void* work, int64_t* iters, double* res, void* stream) {
    if (n <= 0) {
        *iters = 0;
        *res = 0.0;
        return 1;
    }
    CsrOp<double> A{indptr, indices, data};
    return bicgstab_run<CsrOp<double>, double>(n, A, b, inv_diag, x, tol, maxiter, relative_to_initial, work, iters,
                                               res, (hipStream_t)stream, false);
}

int ftmi_bicgstab_ldu_sliced(int64_t n, const double* diag, const int32_t* base, const uint8_t* deg,
                             const int32_t* nbr, const double* off, const double* b, const double* inv_diag, double* x,
                             double tol, int64_t maxiter, int relative_to_initial, void* work, int64_t* iters,
                             double* res, void* stream) {
    if (n <= 0) {
        *iters = 0;
        *res = 0.0;
        return 1;
    }
    LduSellOp<double> A{diag, base, deg, nbr, off};
    return bicgstab_run<LduSellOp<double>, double>(n, A, b, inv_diag, x, tol, maxiter, relative_to_initial, work,
                                                   iters, res, (hipStream_t)stream, false);
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
    return bicgstab_run<LduOp<double>, double>(n, A, b, inv_diag, x, tol, maxiter, relative_to_initial, work, iters,
                                               res, (hipStream_t)stream, false);
}

}  // extern "C"

// ====================================================================== AMG V-cycle
namespace {

template <typename T>
struct Level {
    int64_t n;
    CsrOp<T> A, P, R;
    int tpr_a, tpr_p, tpr_r;  // threads per row of the level's CSR kernels
    int64_t nnz_a, nnz_p;
    const T* Dinv;
    T *b, *x, *y, *r;  // level vectors (level 0: b is the caller's rhs, x its output)
};

// A hierarchy handle holds FP64 or FP32 values; `bits` says which (checked at every entry).
struct AmgBase {
    int bits = 64;
};

template <typename T>
struct Amg : AmgBase {
    int nlev = 0;
    Level<T> lv[24];
    int64_t nc = 0;
    const T* cinv = nullptr;
    T *cb = nullptr, *cx = nullptr;
    // iterative coarse solve (csweeps > 0, instead of the dense inverse): csweeps weighted-
    // Jacobi sweeps from zero on the coarsest matrix cA (a hierarchy cut at a level count)
    int csweeps = 0;
    double comega = 0.0;  // the coarse sweeps' weight (below 2 / rho(D^-1 A_c), so they converge)
    CsrOp<T> cA{};
    const T* cdinv = nullptr;
    T* cy = nullptr;
    int tpr_c = 1;
    int64_t nnz_c = 0;
    double omega = 0.8;
    int sweeps = 2;
    Amg() { bits = 8 * (int)sizeof(T); }
};

// Rows of the coarse Galerkin levels grow long (hundreds of entries on the last levels), so
// each CSR kernel of the cycle gives TPR lanes to a row and reduces them with width-TPR
// shuffles; TPR = 1 keeps the plain column-order sum of the fine levels.
// g_tpr_bias (ftmi_set_tpr_bias, the autotuner's knob) shifts every choice by that many steps
// of 1 / 4 / 16 / 64; it is read when a level's choice is made (add_level, ftmi_amg_retune).
static int g_tpr_bias = 0;

inline int pick_tpr(int64_t n, int64_t nnz) {
    const double avg = n > 0 ? (double)nnz / (double)n : 0.0;
    int k = avg <= 8.0 ? 0 : avg <= 24.0 ? 1 : avg <= 96.0 ? 2 : 3;
    k += g_tpr_bias;
    k = k < 0 ? 0 : k > 3 ? 3 : k;
    return 1 << (2 * k);
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

template <int TPR, typename T>
__device__ inline T seg_sum(T v) {
#pragma unroll
    for (int off = TPR / 2; off > 0; off >>= 1) v += __shfl_xor(v, off, TPR);
    return v;
}

// sum_j a_ij f(j) over the entries lane, lane + TPR, ... of row i, in that order
template <int TPR, typename T, class F>
__device__ inline T lane_sum(const CsrOp<T>& A, int64_t i, int lane, F f) {
    const int32_t p = A.ptr[i] + lane, e = A.ptr[i + 1];
    const int d = p < e ? (e - p + TPR - 1) / TPR : 0;
    return gather_sum(
        T(0), d, 0, [&](int k) { return A.col[p + TPR * k]; }, [&](int k) { return A.val[p + TPR * k]; }, f);
}

template <int TPR, typename T>
__device__ inline T row_dot(const CsrOp<T>& A, int64_t i, const T* __restrict__ x, int lane) {
    return seg_sum<TPR>(lane_sum<TPR>(A, i, lane, [&](int32_t j) { return x[j]; }));
}

inline int grid_rows(int64_t n, int tpr) { return grid_for(n * tpr); }

// The end of PCG iteration k, folded into the first kernel of its preconditioner: ||r|| from
// the partials of the x/r update, the convergence test and the host post (the work of a
// one-workgroup finalize kernel). `part` == nullptr: nothing to finish.
struct PcgFinish {
    State* S;
    const double* part;
    int np;
    double tol;
    const double* tol_dev;
    int k;
    Post post;
};

// true when the solve is done (then the caller returns); every thread gets the same answer
__device__ inline bool pcg_finish(const PcgFinish& f) {
    const bool lead = blockIdx.x == 0 && threadIdx.x == 0;
    State* S = f.S;
    // the state words and the tolerance are loaded with the done flag, ahead of the partial sums
    const int done0 = S->st[I_DONE];
    const double bnorm = S->sc[S_BNORM];
    const double tol = f.tol_dev ? *f.tol_dev : f.tol;
    if (done0) {
        if (lead) post_host(S, f.post);
        return true;
    }
    double tot[1];
    consume_sum<1>(f.part, f.np, tot);
    const double rn = sqrt(tot[0]);
    const double res = rn / bnorm;
    const bool done = res < tol;
    if (lead) {
        S->sc[S_RNORM] = rn;
        S->sc[S_RES] = res;
        if (done) {
            S->st[I_DONE] = 1;
            S->st[I_CONV] = 1;
            S->st[I_ITERS] = f.k;
        }
        post_host(S, f.post);
    }
    return done;
}

// x = w D b + w D (b - A (w D b)): the first two pre-smoothing sweeps from x = 0, fused.
template <int TPR, typename T>
__global__ void __launch_bounds__(kBlock) vc_pre2(int64_t n, CsrOp<T> A, const T* __restrict__ Dinv,
                                                  const T* __restrict__ b, T* __restrict__ x, double w_d,
                                                  const int* done, PcgFinish fin) {
    if (fin.part) {
        if (pcg_finish(fin)) return;
    } else if (done && *done) {
        return;
    }
    const T w = (T)w_d;
    RowGroups<TPR> G;
    for (int64_t i = G.first; i < n; i += G.stride) {
        T ax = lane_sum<TPR>(A, i, G.lane, [&](int32_t c) { return (w * Dinv[c]) * b[c]; });
        ax = seg_sum<TPR>(ax);
        if (G.lane == 0) {
            const T wd = w * Dinv[i];
            x[i] = wd * b[i] + wd * (b[i] - ax);
        }
    }
}

// x = w D b (first sweep only; used when sweeps == 1)
template <typename T>
__global__ void __launch_bounds__(kBlock) vc_pre1(int64_t n, const T* __restrict__ Dinv, const T* __restrict__ b,
                                                  T* __restrict__ x, double w_d, const int* done) {
    if (done && *done) return;
    const T w = (T)w_d;
    int64_t lo, hi;
    block_range(n, lo, hi);
    for (int64_t i = lo + threadIdx.x; i < hi; i += kBlock) x[i] = (w * Dinv[i]) * b[i];
}

// y = x + w D (b - A x), optionally with the workgroup partial sums of y (RED = kRedSum) or
// of b . y (RED = kRedDot: with b the PCG residual r and y = M r, the r . z of the iteration)
enum { kRedNone = 0, kRedSum = 1, kRedDot = 2 };
template <int TPR, int RED, typename T>
__global__ void __launch_bounds__(kBlock) vc_jacobi(int64_t n, CsrOp<T> A, const T* __restrict__ Dinv,
                                                    const T* __restrict__ b, const T* __restrict__ x,
                                                    T* __restrict__ y, double w_d, const int* done,
                                                    double* sum_partials) {
    if (done && *done) return;
    const T w = (T)w_d;
    RowGroups<TPR> G;
    T acc[1] = {T(0)};
    for (int64_t i = G.first; i < n; i += G.stride) {
        const T ax = row_dot<TPR>(A, i, x, G.lane);
        if (G.lane == 0) {
            const T bi = b[i];
            const T yi = x[i] + (w * Dinv[i]) * (bi - ax);
            y[i] = yi;
            if (RED == kRedSum) acc[0] += yi;
            if (RED == kRedDot) acc[0] += bi * yi;
        }
    }
    if (RED != kRedNone) block_partials<1>(acc, sum_partials);
}

template <int TPR, typename T>
__global__ void __launch_bounds__(kBlock) vc_resid(int64_t n, CsrOp<T> A, const T* __restrict__ b,
                                                   const T* __restrict__ x, T* __restrict__ r, const int* done) {
    if (done && *done) return;
    RowGroups<TPR> G;
    for (int64_t i = G.first; i < n; i += G.stride) {
        const T ax = row_dot<TPR>(A, i, x, G.lane);
        if (G.lane == 0) r[i] = b[i] - ax;
    }
}

template <int TPR, typename T>
__global__ void __launch_bounds__(kBlock) vc_spmv(int64_t n, CsrOp<T> A, const T* __restrict__ x, T* __restrict__ y,
                                                  const int* done) {
    if (done && *done) return;
    RowGroups<TPR> G;
    for (int64_t i = G.first; i < n; i += G.stride) {
        const T ax = row_dot<TPR>(A, i, x, G.lane);
        if (G.lane == 0) y[i] = ax;
    }
}

template <int TPR, typename T>
__global__ void __launch_bounds__(kBlock) vc_prolong(int64_t n, CsrOp<T> P, const T* __restrict__ e,
                                                     T* __restrict__ x, const int* done) {
    if (done && *done) return;
    RowGroups<TPR> G;
    for (int64_t i = G.first; i < n; i += G.stride) {
        const T pe = row_dot<TPR>(P, i, e, G.lane);
        if (G.lane == 0) x[i] = x[i] + pe;
    }
}

template <typename T>
__global__ void __launch_bounds__(kBlock) vc_copy(int64_t n, const T* __restrict__ a, T* __restrict__ b,
                                                  const int* done) {
    if (done && *done) return;
    int64_t lo, hi;
    block_range(n, lo, hi);
    for (int64_t i = lo + threadIdx.x; i < hi; i += kBlock) b[i] = a[i];
}

// workgroup partial sums of z (the PCG consumer turns them into the mean)
template <typename T>
__global__ void __launch_bounds__(kBlock) k_sum(int64_t n, const T* __restrict__ z, const int* done,
                                                double* sum_partials) {
    if (done && *done) return;
    int64_t lo, hi;
    block_range(n, lo, hi);
    T acc[1] = {T(0)};
    for (int64_t i = lo + threadIdx.x; i < hi; i += kBlock) acc[0] += z[i];
    block_partials<1>(acc, sum_partials);
}

// x = Minv b for the dense coarse inverse (row per wave)
template <typename T>
__global__ void __launch_bounds__(kBlock) vc_coarse(int64_t n, const T* __restrict__ M, const T* __restrict__ b,
                                                    T* __restrict__ x, const int* done) {
    if (done && *done) return;
    const int lane = threadIdx.x % kWave;
    const int64_t row = (int64_t)blockIdx.x * (kBlock / kWave) + threadIdx.x / kWave;
    if (row >= n) return;
    T s = T(0);
    for (int64_t j = lane; j < n; j += kWave) s += M[row * n + j] * b[j];
    s = wave_sum(s);
    if (lane == 0) x[row] = s;
}

#define TPR_LAUNCH(tpr, KERNEL, n, st, ...)                                                              \
    do {                                                                                                 \
        switch (tpr) {                                                                                   \
            case 1: KERNEL<1, T><<<grid_rows(n, 1), kBlock, 0, st>>>(__VA_ARGS__); break;                 \
            case 4: KERNEL<4, T><<<grid_rows(n, 4), kBlock, 0, st>>>(__VA_ARGS__); break;                 \
            case 16: KERNEL<16, T><<<grid_rows(n, 16), kBlock, 0, st>>>(__VA_ARGS__); break;              \
            default: KERNEL<64, T><<<grid_rows(n, 64), kBlock, 0, st>>>(__VA_ARGS__); break;              \
        }                                                                                                \
    } while (0)

// grid of the summing smoother: its partials are consumed by every workgroup of the next kernel
inline int sum_grid(int64_t n, int tpr) {
    const int g = grid_rows(n, tpr);
    return g < kRedGrid ? g : kRedGrid;
}

template <int TPR, typename T>
static void launch_jacobi(int64_t n, CsrOp<T> A, const T* Dinv, const T* b, const T* x, T* y, double w,
                          const int* done, double* part, int red, hipStream_t st) {
    if (part && red == kRedSum)
        vc_jacobi<TPR, kRedSum, T><<<sum_grid(n, TPR), kBlock, 0, st>>>(n, A, Dinv, b, x, y, w, done, part);
    else if (part && red == kRedDot)
        vc_jacobi<TPR, kRedDot, T><<<sum_grid(n, TPR), kBlock, 0, st>>>(n, A, Dinv, b, x, y, w, done, part);
    else
        vc_jacobi<TPR, kRedNone, T><<<grid_rows(n, TPR), kBlock, 0, st>>>(n, A, Dinv, b, x, y, w, done, nullptr);
}

template <typename T>
static void jacobi(int tpr, int64_t n, CsrOp<T> A, const T* Dinv, const T* b, const T* x, T* y, double w,
                   const int* done, double* part, int red, hipStream_t st) {
    switch (tpr) {
        case 1: launch_jacobi<1, T>(n, A, Dinv, b, x, y, w, done, part, red, st); break;
        case 4: launch_jacobi<4, T>(n, A, Dinv, b, x, y, w, done, part, red, st); break;
        case 16: launch_jacobi<16, T>(n, A, Dinv, b, x, y, w, done, part, red, st); break;
        default: launch_jacobi<64, T>(n, A, Dinv, b, x, y, w, done, part, red, st); break;
    }
}

constexpr PcgFinish kNoFinish{nullptr, nullptr, 0, 0.0, nullptr, 0, kNoPost};

// Enqueue one V-cycle on level k: x_k = M_k^{-1} b_k. With `part`, the last smoothing sweep
// also writes workgroup partial sums (sum_grid of them): of the output (red = kRedSum) or of
// b_k . x_k (red = kRedDot). With fin.part set, the level's first kernel first finishes the PCG
// iteration that produced b_k (pcg_finish); it needs at least two sweeps.
template <typename T>
static void vcycle(const Amg<T>& M, int k, hipStream_t st, const int* done, double* part, int red = kRedSum,
                   PcgFinish fin = kNoFinish) {
    if (k == M.nlev && M.csweeps >= 2) {
        // weighted-Jacobi sweeps from zero: a symmetric positive definite approximate inverse
        // (a polynomial in D^-1 A times D^-1), so the cycle stays a valid PCG preconditioner
        TPR_LAUNCH(M.tpr_c, vc_pre2, M.nc, st, M.nc, M.cA, M.cdinv, M.cb, M.cx, M.comega, done, kNoFinish);
        T* cur = M.cx;
        T* other = M.cy;
        for (int s = 2; s < M.csweeps; ++s) {
            jacobi<T>(M.tpr_c, M.nc, M.cA, M.cdinv, M.cb, cur, other, M.comega, done, nullptr, kRedNone, st);
            T* t = cur;
            cur = other;
            other = t;
        }
        if (cur != M.cx) vc_copy<T><<<grid_for(M.nc), kBlock, 0, st>>>(M.nc, cur, M.cx, done);
        return;
    }
    if (k == M.nlev) {
        const int rows_per_block = kBlock / kWave;
        vc_coarse<T><<<(int)((M.nc + rows_per_block - 1) / rows_per_block), kBlock, 0, st>>>(M.nc, M.cinv, M.cb,
                                                                                             M.cx, done);
        return;
    }
    const Level<T>& L = M.lv[k];
    const double w = M.omega;
    T* next_b = (k + 1 == M.nlev) ? M.cb : M.lv[k + 1].b;
    T* next_x = (k + 1 == M.nlev) ? M.cx : M.lv[k + 1].x;
    const int64_t next_n = (k + 1 == M.nlev) ? M.nc : M.lv[k + 1].n;
    // pre-smoothing: `sweeps` weighted-Jacobi sweeps from zero
    T* xa = L.x;
    T* xb = L.y;
    if (M.sweeps >= 2) {
        TPR_LAUNCH(L.tpr_a, vc_pre2, L.n, st, L.n, L.A, L.Dinv, L.b, xa, w, done, fin);
        for (int s = 2; s < M.sweeps; ++s) {
            jacobi<T>(L.tpr_a, L.n, L.A, L.Dinv, L.b, xa, xb, w, done, nullptr, kRedNone, st);
            T* t = xa;
            xa = xb;
            xb = t;
        }
    } else {
        vc_pre1<T><<<grid_for(L.n), kBlock, 0, st>>>(L.n, L.Dinv, L.b, xa, w, done);
    }
    TPR_LAUNCH(L.tpr_a, vc_resid, L.n, st, L.n, L.A, L.b, xa, L.r, done);
    TPR_LAUNCH(L.tpr_r, vc_spmv, next_n, st, next_n, L.R, L.r, next_b, done);
    vcycle(M, k + 1, st, done, nullptr);
    TPR_LAUNCH(L.tpr_p, vc_prolong, L.n, st, L.n, L.P, next_x, xa, done);
    // post-smoothing: ping-pong between the level's two buffers, never in place
    T* cur = xa;
    T* other = (xa == L.x) ? L.y : L.x;
    for (int s = 0; s < M.sweeps; ++s) {
        const bool last = (s == M.sweeps - 1);
        jacobi<T>(L.tpr_a, L.n, L.A, L.Dinv, L.b, cur, other, w, done, last ? part : nullptr, red, st);
        T* t = cur;
        cur = other;
        other = t;
    }
    if (cur != L.x) vc_copy<T><<<grid_for(L.n), kBlock, 0, st>>>(L.n, cur, L.x, done);
}

}  // namespace


// ------------------------------------------------------------------ hierarchy set-up (host)
namespace {

template <typename T>
Amg<T>* amg_of(void* h, const char* where) {
    AmgBase* b = (AmgBase*)h;
    if (!b || b->bits != 8 * (int)sizeof(T)) {
        set_error_msg(where);
        return nullptr;
    }
    return (Amg<T>*)b;
}

template <typename T>
int amg_add_level(void* h, int64_t n, int64_t nnz_a, int64_t nnz_p, const int32_t* a_ptr, const int32_t* a_col,
                  const T* a_val,
                  const int32_t* p_ptr, const int32_t* p_col, const T* p_val, const int32_t* r_ptr,
                  const int32_t* r_col, const T* r_val, const T* dinv, T* vb, T* vx, T* vy, T* vr) {
    Amg<T>* M = amg_of<T>(h, "ftmi_amg_add_level: handle precision does not match the values");
    if (!M) return -1;
    if (M->nlev >= 24) {
        set_error_msg("ftmi_amg_add_level: at most 24 levels");
        return -1;
    }
    Level<T>& L = M->lv[M->nlev++];
    L.n = n;
    L.A = CsrOp<T>{a_ptr, a_col, a_val};
    L.P = CsrOp<T>{p_ptr, p_col, p_val};
    L.R = CsrOp<T>{r_ptr, r_col, r_val};
    L.tpr_a = pick_tpr(n, nnz_a);
    L.tpr_p = pick_tpr(n, nnz_p);
    L.nnz_a = nnz_a;
    L.nnz_p = nnz_p;
    L.tpr_r = 1;  // R = P^T has the next level's rows: set when that size is known
    if (M->nlev >= 2) M->lv[M->nlev - 2].tpr_r = pick_tpr(n, M->lv[M->nlev - 2].nnz_p);
    L.Dinv = dinv;
    L.b = vb;
    L.x = vx;
    L.y = vy;
    L.r = vr;
    return 0;
}

template <typename T>
int amg_set_coarse(void* h, int64_t n, const T* inv, T* vb, T* vx) {
    Amg<T>* M = amg_of<T>(h, "ftmi_amg_set_coarse: handle precision does not match the values");
    if (!M) return -1;
    M->nc = n;
    if (M->nlev >= 1) M->lv[M->nlev - 1].tpr_r = pick_tpr(n, M->lv[M->nlev - 1].nnz_p);
    M->cinv = inv;
    M->cb = vb;
    M->cx = vx;
    return 0;
}

template <typename T>
int amg_set_coarse_iter(void* h, int64_t n, int64_t nnz, const int32_t* a_ptr, const int32_t* a_col,
                        const T* a_val, const T* dinv, int sweeps, double omega, T* vb, T* vx, T* vy) {
    Amg<T>* M = amg_of<T>(h, "ftmi_amg_set_coarse_iter: handle precision does not match the values");
    if (!M) return -1;
    if (sweeps < 2 || M->nlev < 1) {
        set_error_msg("ftmi_amg_set_coarse_iter: needs at least two sweeps and one level above the coarse one");
        return -1;
    }
    M->nc = n;
    M->lv[M->nlev - 1].tpr_r = pick_tpr(n, M->lv[M->nlev - 1].nnz_p);
    M->cinv = nullptr;
    M->csweeps = sweeps;
    M->comega = omega;
    M->cA = CsrOp<T>{a_ptr, a_col, a_val};
    M->nnz_c = nnz;
    M->tpr_c = pick_tpr(n, nnz);
    M->cdinv = dinv;
    M->cb = vb;
    M->cx = vx;
    M->cy = vy;
    return 0;
}

// the threads-per-row choices of every level again, with the current bias
template <typename T>
void amg_retune(Amg<T>* M) {
    for (int l = 0; l < M->nlev; ++l) {
        Level<T>& L = M->lv[l];
        L.tpr_a = pick_tpr(L.n, L.nnz_a);
        L.tpr_p = pick_tpr(L.n, L.nnz_p);
        const int64_t rows_r = l + 1 < M->nlev ? M->lv[l + 1].n : M->nc;
        L.tpr_r = pick_tpr(rows_r, L.nnz_p);
    }
    if (M->csweeps > 0) M->tpr_c = pick_tpr(M->nc, M->nnz_c);
}

template <typename T>
int amg_vcycle(void* h, const T* b, T* x, void* stream) {
    Amg<T>* M = amg_of<T>(h, "ftmi_amg_vcycle: handle precision does not match the vectors");
    if (!M) return -1;
    hipStream_t st = (hipStream_t)stream;
    if (M->nlev == 0) {
        const int rows_per_block = kBlock / kWave;
        vc_coarse<T><<<(int)((M->nc + rows_per_block - 1) / rows_per_block), kBlock, 0, st>>>(M->nc, M->cinv, b, x,
                                                                                              nullptr);
        FTMI_LAUNCH_CHECK("ftmi_amg_vcycle");
        return 0;
    }
    Level<T>& L0 = M->lv[0];
    T *sb = L0.b, *sx = L0.x;
    L0.b = (T*)b;
    L0.x = x;
    vcycle<T>(*M, 0, st, nullptr, nullptr);
    L0.b = sb;
    L0.x = sx;
    FTMI_LAUNCH_CHECK("ftmi_amg_vcycle");
    return 0;
}

}  // namespace

extern "C" {

int ftmi_set_tpr_bias(int bias) {
    const int was = g_tpr_bias;
    g_tpr_bias = bias < -3 ? -3 : bias > 3 ? 3 : bias;
    return was;
}

int ftmi_amg_retune(void* h) {
    AmgBase* b = (AmgBase*)h;
    if (!b) {
        set_error_msg("ftmi_amg_retune: null hierarchy");
        return -1;
    }
    if (b->bits == 64)
        amg_retune((Amg<double>*)b);
    else
        amg_retune((Amg<float>*)b);
    return 0;
}


void* ftmi_amg_create(double omega, int sweeps) {
    Amg<double>* M = new Amg<double>();
    M->omega = omega;
    M->sweeps = sweeps;
    return M;
}

void* ftmi_amg_create_f32(double omega, int sweeps) {
    Amg<float>* M = new Amg<float>();
    M->omega = omega;
    M->sweeps = sweeps;
    return M;
}

void ftmi_amg_destroy(void* h) {
    AmgBase* b = (AmgBase*)h;
    if (!b) return;
    if (b->bits == 32)
        delete (Amg<float>*)b;
    else
        delete (Amg<double>*)b;
}

int ftmi_amg_add_level(void* h, int64_t n, int64_t nnz_a, int64_t nnz_p, const int32_t* a_ptr, const int32_t* a_col,
                       const double* a_val, const int32_t* p_ptr, const int32_t* p_col, const double* p_val,
                       const int32_t* r_ptr, const int32_t* r_col, const double* r_val, const double* dinv, double* vb,
                       double* vx, double* vy, double* vr) {
    return amg_add_level<double>(h, n, nnz_a, nnz_p, a_ptr, a_col, a_val, p_ptr, p_col, p_val, r_ptr, r_col, r_val, dinv, vb, vx,
                                 vy, vr);
}

int ftmi_amg_add_level_f32(void* h, int64_t n, int64_t nnz_a, int64_t nnz_p, const int32_t* a_ptr,
                           const int32_t* a_col, const float* a_val, const int32_t* p_ptr, const int32_t* p_col,
                           const float* p_val, const int32_t* r_ptr, const int32_t* r_col, const float* r_val,
                           const float* dinv, float* vb, float* vx, float* vy, float* vr) {
    return amg_add_level<float>(h, n, nnz_a, nnz_p, a_ptr, a_col, a_val, p_ptr, p_col, p_val, r_ptr, r_col, r_val, dinv, vb, vx,
                                vy, vr);
}

int ftmi_amg_set_coarse(void* h, int64_t n, const double* inv, double* vb, double* vx) {
    return amg_set_coarse<double>(h, n, inv, vb, vx);
}

int ftmi_amg_set_coarse_f32(void* h, int64_t n, const float* inv, float* vb, float* vx) {
    return amg_set_coarse<float>(h, n, inv, vb, vx);
}

int ftmi_amg_set_coarse_iter(void* h, int64_t n, int64_t nnz, const int32_t* a_ptr, const int32_t* a_col,
                             const double* a_val, const double* dinv, int sweeps, double omega, double* vb,
                             double* vx, double* vy) {
    return amg_set_coarse_iter<double>(h, n, nnz, a_ptr, a_col, a_val, dinv, sweeps, omega, vb, vx, vy);
}

int ftmi_amg_set_coarse_iter_f32(void* h, int64_t n, int64_t nnz, const int32_t* a_ptr, const int32_t* a_col,
                                 const float* a_val, const float* dinv, int sweeps, double omega, float* vb,
                                 float* vx, float* vy) {
    return amg_set_coarse_iter<float>(h, n, nnz, a_ptr, a_col, a_val, dinv, sweeps, omega, vb, vx, vy);
}

// x = M^{-1} b with one V-cycle (b, x: level-0 vectors, distinct from the level buffers).
int ftmi_amg_vcycle(void* h, const double* b, double* x, void* stream) { return amg_vcycle<double>(h, b, x, stream); }

int ftmi_amg_vcycle_f32(void* h, const float* b, float* x, void* stream) { return amg_vcycle<float>(h, b, x, stream); }

}  // extern "C"

// ====================================================================== PCG
// One iteration: q = A p with the p update fused in (p = z + beta p_prev, recomputed at the
// neighbours), the x/r update, the state finalize the host polls, the V-cycle, and the
// projection of z with r.z. p alternates between two buffers.
namespace {

template <bool PROJ, typename T>
__global__ void __launch_bounds__(kBlock) pcg_init_a(int64_t n, const T* __restrict__ b, double* part_out) {
    int64_t lo, hi;
    block_range(n, lo, hi);
    T acc[2] = {T(0), T(0)};
    for (int64_t i = lo + threadIdx.x; i < hi; i += kBlock) {
        acc[0] += b[i];
        acc[1] += b[i] * b[i];
    }
    block_partials<2>(acc, part_out);
}

// bnorm, mean(b); r = b - mean (projection) or b; x = 0 unless the solve adds to x (then it
// solves for the correction of the given x, whose residual b is); partials r.r
template <bool PROJ, typename T, typename XT>
__global__ void __launch_bounds__(kBlock) pcg_init_b(int64_t n, const T* __restrict__ b, T* __restrict__ r,
                                                     XT* __restrict__ x, int zero_x, State* S, const double* part_in,
                                                     int np_in, double* part_out) {
    double tot[2];
    consume_sum<2>(part_in, np_in, tot);
    const double mean_d = PROJ ? tot[0] / (double)n : 0.0;
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        const double bn = sqrt(tot[1]);
        S->sc[S_BNORM] = bn != 0.0 ? bn : 1.0;
        S->sc[S_BMEAN] = mean_d;
    }
    const T mean = (T)mean_d;
    int64_t lo, hi;
    block_range(n, lo, hi);
    T acc[1] = {T(0)};
    for (int64_t i = lo + threadIdx.x; i < hi; i += kBlock) {
        const T ri = PROJ ? b[i] - mean : b[i];
        r[i] = ri;
        if (zero_x) x[i] = XT(0);
        acc[0] += ri * ri;
    }
    block_partials<1>(acc, part_out);
}

__global__ void __launch_bounds__(kBlock) fin_pcg_init(const double* partials, int nparts, State* S, double tol,
                                                       Post post, const double* tol_dev) {
    double tot[1];
    if (tol_dev) tol = *tol_dev;
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
    post_host(S, post);
}

// z -= mean(z) (mean from the V-cycle's partial sums); partials r.z
template <bool PROJ, typename T>
__global__ void __launch_bounds__(kBlock) pcg_z(int64_t n, const T* __restrict__ r, T* __restrict__ z, const State* S,
                                                const double* zsum_in, int np_zsum, double* part_out) {
    if (S->st[I_DONE]) return;
    double mean_d = 0.0;
    if (PROJ) {
        double tot[1];
        consume_sum<1>(zsum_in, np_zsum, tot);
        mean_d = tot[0] / (double)n;
    }
    const T mean = (T)mean_d;
    int64_t lo, hi;
    block_range(n, lo, hi);
    T acc[1] = {T(0)};
    for (int64_t i = lo + threadIdx.x; i < hi; i += kBlock) {
        T zi = z[i];
        if (PROJ) {
            zi = zi - mean;
            z[i] = zi;
        }
        acc[0] += r[i] * zi;
    }
    block_partials<1>(acc, part_out);
}

// rz_new from the partials; FIRST: p = z, rz = rz_new; else breakdown if rz == 0, beta =
// rz_new / rz, p = z + beta p_prev; then q = A p and partials p.q. FIRST is a template
// argument: as a run-time flag it put a branch, and a wait for its load, inside every
// neighbour gather, so the gathers of a row no longer overlapped.
template <class Op, typename T, bool FIRST>
__global__ void __launch_bounds__(kBlock) pcg_ap(int64_t n, Op A, const T* __restrict__ z, const T* __restrict__ p_prev,
                                                 T* __restrict__ p, T* __restrict__ q, State* S, const double* part_in,
                                                 int np_in, double* part_out, int k) {
    const int cur = S_RZ + (k & 1), prev = S_RZ + ((k - 1) & 1);
    const int done = S->st[I_DONE];
    const double rz = FIRST ? 0.0 : S->sc[prev];
    if (done) return;
    double tot[1];
    consume_sum<1>(part_in, np_in, tot);
    const bool lead = blockIdx.x == 0 && threadIdx.x == 0;
    double beta_d = 0.0;
    if (!FIRST) {
        if (rz == 0.0) {  // breakdown found at the end of the previous iteration
            if (lead) {
                S->st[I_DONE] = 1;
                S->st[I_CONV] = 0;
                S->st[I_ITERS] = k - 1;
            }
            return;
        }
        beta_d = tot[0] / rz;
    }
    if (lead) S->sc[cur] = tot[0];
    const T beta = (T)beta_d;
    auto pf = [&](int64_t j) {
        if (FIRST) return z[j];
        return z[j] + beta * p_prev[j];
    };
    int64_t lo, hi;
    block_range(n, lo, hi);
    T acc[1] = {T(0)};
    for (int64_t i = lo + threadIdx.x; i < hi; i += kBlock) {
        const T pi = pf(i);
        p[i] = pi;
        const T qi = A.row_f(i, pf);
        q[i] = qi;
        acc[0] += pi * qi;
    }
    block_partials<1>(acc, part_out);
}

// alpha = rz / (p.q) (breakdown if p.q <= 0); x += alpha p (in the iterate's precision XT);
// r -= alpha q; partials r.r
template <typename T, typename XT>
__global__ void __launch_bounds__(kBlock) pcg_xr(int64_t n, const T* __restrict__ p, const T* __restrict__ q,
                                                 XT* __restrict__ x, T* __restrict__ r, State* S,
                                                 const double* part_in, int np_in, double* part_out, int k) {
    const int done = S->st[I_DONE];
    const double rz = S->sc[S_RZ + (k & 1)];
    if (done) return;
    double tot[1];
    consume_sum<1>(part_in, np_in, tot);
    const bool lead = blockIdx.x == 0 && threadIdx.x == 0;
    if (tot[0] <= 0.0) {
        if (lead) {
            S->st[I_DONE] = 1;
            S->st[I_CONV] = 0;
            S->st[I_ITERS] = k;
        }
        return;
    }
    const double a_d = rz / tot[0];
    const T a = (T)a_d;
    const XT ax = (XT)a_d;
    int64_t lo, hi;
    block_range(n, lo, hi);
    T acc[1] = {T(0)};
    for (int64_t i = lo + threadIdx.x; i < hi; i += kBlock) {
        x[i] += ax * (XT)p[i];
        const T ri = r[i] - a * q[i];
        r[i] = ri;
        acc[0] += ri * ri;
    }
    block_partials<1>(acc, part_out);
}

__global__ void __launch_bounds__(kBlock) fin_pcg_xr(const double* partials, int nparts, State* S, double tol, int k,
                                                     Post post, const double* tol_dev) {
    double tot[1];
    if (tol_dev) tol = *tol_dev;
    if (!S->st[I_DONE] && finalize_sum<1>(partials, nparts, tot)) {
        const double rn = sqrt(tot[0]);
        S->sc[S_RNORM] = rn;
        S->sc[S_RES] = rn / S->sc[S_BNORM];
        if (rn / S->sc[S_BNORM] < tol) {
            S->st[I_DONE] = 1;
            S->st[I_CONV] = 1;
            S->st[I_ITERS] = k;
        }
    }
    post_host(S, post);
}

template <typename T>
struct PcgWork {
    T *r, *z, *p[2], *q;
    State* S;
    unsigned* counter;
    double* part[3];
};

template <typename T>
static int64_t pcg_bytes(int64_t n) {
    return align256(5 * n * (int64_t)sizeof(T)) + 512 + 3 * kPartBytes;
}

template <typename T>
static PcgWork<T> carve_pcg(int64_t n, void* work) {
    char* w = (char*)work;
    PcgWork<T> W;
    T* d = (T*)w;
    W.r = d;
    W.z = d + n;
    W.p[0] = d + 2 * n;
    W.p[1] = d + 3 * n;
    W.q = d + 4 * n;
    char* tail = w + align256(5 * n * (int64_t)sizeof(T));
    W.S = (State*)tail;
    W.counter = (unsigned*)(tail + 256);
    for (int j = 0; j < 3; ++j) W.part[j] = (double*)(tail + 512 + j * kPartBytes);
    return W;
}

// x in the iterate precision XT; zero_x = false adds the solve's result to the given x
template <class Op, typename T, typename XT = T>
int pcg_run(int64_t n, Op A, const T* b, XT* x, double tol, int64_t maxiter, Amg<T>* M, int project, void* work,
            int64_t* iters, double* res, hipStream_t st, bool zero_x = true, const double* tol_dev = nullptr,
            int64_t cap = -1, double* report = nullptr) {
    Pinned* pn = pinned();
    if (!pn) {
        set_error_msg("pcg: pinned host memory unavailable");
        return -1;
    }
    PcgWork<T> W = carve_pcg<T>(n, work);
    const int g = grid_red(n);
    const int* done = &W.S->st[I_DONE];
    double *P0 = W.part[0], *P1 = W.part[1], *P2 = W.part[2];
    k_state_reset<<<1, 64, 0, st>>>(W.S, W.counter);
    const int zx = zero_x ? 1 : 0;
    if (project) {
        pcg_init_a<true, T><<<g, kBlock, 0, st>>>(n, b, P0);
        pcg_init_b<true, T, XT><<<g, kBlock, 0, st>>>(n, b, W.r, x, zx, W.S, P0, g, P1);
    } else {
        pcg_init_a<false, T><<<g, kBlock, 0, st>>>(n, b, P0);
        pcg_init_b<false, T, XT><<<g, kBlock, 0, st>>>(n, b, W.r, x, zx, W.S, P0, g, P1);
    }
    fin_pcg_init<<<1, kBlock, 0, st>>>(P1, g, W.S, tol, g_capture ? kNoPost : next_post(pn, 0), tol_dev);
    // z = M r and the partials of r . z in P0; returns their count. With a V-cycle of two or
    // more sweeps, `fin` (the end of the previous iteration: ||r||, convergence, host post) is
    // done by the cycle's first kernel and r . z by its last sweep (with the projection, the
    // last sweep sums z and pcg_z removes the mean and forms r . z); otherwise `fin` is a
    // finalize kernel of its own and pcg_z forms r . z.
    const bool fold = M && M->nlev > 0 && M->sweeps >= 2;
    auto precond = [&](PcgFinish fin) -> int {
        if (fold) {
            Level<T>& L0 = M->lv[0];
            T *sb = L0.b, *sx = L0.x;
            L0.b = W.r;
            L0.x = W.z;
            const int nz = sum_grid(n, L0.tpr_a);
            vcycle<T>(*M, 0, st, done, project ? P2 : P0, project ? kRedSum : kRedDot, fin);
            L0.b = sb;
            L0.x = sx;
            if (!project) return nz;
            pcg_z<true, T><<<g, kBlock, 0, st>>>(n, W.r, W.z, W.S, P2, nz, P0);
            return g;
        }
        if (fin.part) fin_pcg_xr<<<1, kBlock, 0, st>>>(fin.part, fin.np, W.S, fin.tol, fin.k, fin.post, fin.tol_dev);
        int nz = g;
        if (M && M->nlev > 0) {
            Level<T>& L0 = M->lv[0];
            T *sb = L0.b, *sx = L0.x;
            L0.b = W.r;
            L0.x = W.z;
            vcycle<T>(*M, 0, st, done, project ? P2 : nullptr);
            L0.b = sb;
            L0.x = sx;
            nz = sum_grid(n, L0.tpr_a);
        } else {
            if (M) {
                const int rpb = kBlock / kWave;
                vc_coarse<T><<<(int)((M->nc + rpb - 1) / rpb), kBlock, 0, st>>>(M->nc, M->cinv, W.r, W.z, done);
            } else {
                vc_copy<T><<<grid_for(n), kBlock, 0, st>>>(n, W.r, W.z, done);
            }
            if (project) k_sum<T><<<g, kBlock, 0, st>>>(n, W.z, done, P2);
        }
        if (project)
            pcg_z<true, T><<<g, kBlock, 0, st>>>(n, W.r, W.z, W.S, P2, nz, P0);
        else
            pcg_z<false, T><<<g, kBlock, 0, st>>>(n, W.r, W.z, W.S, P2, nz, P0);
        return g;
    };
    int nrz = precond(kNoFinish);
    FTMI_LAUNCH_CHECK("pcg init");
    auto launch_ap = [&](int64_t k, int a, int c) {
        if (k == 1)
            pcg_ap<Op, T, true><<<g, kBlock, 0, st>>>(n, A, W.z, W.p[a], W.p[c], W.q, W.S, P0, nrz, P1, (int)k);
        else
            pcg_ap<Op, T, false><<<g, kBlock, 0, st>>>(n, A, W.z, W.p[a], W.p[c], W.q, W.S, P0, nrz, P1, (int)k);
    };
    if (g_capture) {
        if (cap < 0) report = capture_slot(maxiter, &cap);
        for (int64_t k = 1; k <= cap; ++k) {
            const int a = (int)((k - 1) & 1), c = (int)(k & 1);
            launch_ap(k, a, c);
            pcg_xr<T, XT><<<g, kBlock, 0, st>>>(n, W.p[c], W.q, x, W.r, W.S, P1, g, P2, (int)k);
            nrz = precond(PcgFinish{W.S, P2, g, tol, tol_dev, (int)k, kNoPost});
        }
        if (report) k_report<<<1, 64, 0, st>>>(W.S, report, cap >= maxiter ? 1 : 0, cap);
        FTMI_LAUNCH_CHECK("pcg (capture)");
        *iters = -1;
        *res = __builtin_nan("");
        return 0;
    }
    int last = 0;
    for (int64_t k = 1; k <= maxiter; ++k) {
        const int a = (int)((k - 1) & 1), c = (int)(k & 1);
        launch_ap(k, a, c);
        pcg_xr<T, XT><<<g, kBlock, 0, st>>>(n, W.p[c], W.q, x, W.r, W.S, P1, g, P2, (int)k);
        nrz = precond(PcgFinish{W.S, P2, g, tol, tol_dev, (int)k, next_post(pn, last ^ 1)});
        FTMI_LAUNCH_CHECK("pcg iteration");
        if (wait_state(pn, last, st)) return -1;
        const bool prev_done = pn->h[last].st[I_DONE] != 0;
        last ^= 1;
        if (prev_done) break;
    }
    if (wait_state(pn, last, st)) return -1;
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

// ------------------------------------------------------------------ mixed precision
// FP32 right-hand side of a correction solve: rb = b - A x with FP32 matrix and arithmetic
// and the FP64 iterate rounded on load
__global__ void __launch_bounds__(kBlock) k_resid_f32(int64_t n, LduDiffOp<float> A, const float* __restrict__ b,
                                                      const double* __restrict__ x, float* __restrict__ rb) {
    int64_t lo, hi;
    block_range(n, lo, hi);
    auto xf = [&](int64_t j) { return (float)x[j]; };
    for (int64_t i = lo + threadIdx.x; i < hi; i += kBlock) rb[i] = b[i] - A.row_f(i, xf);
}

// x += d (FP32 correction added to the FP64 iterate)
__global__ void __launch_bounds__(kBlock) k_add_f32(int64_t n, const float* __restrict__ d, double* __restrict__ x) {
    int64_t lo, hi;
    block_range(n, lo, hi);
    for (int64_t i = lo + threadIdx.x; i < hi; i += kBlock) x[i] += (double)d[i];
}

// FP64 residual of the pressure system: r = b - A x; partials sum r, sum r^2, sum b^2
__global__ void __launch_bounds__(kBlock) k_resid_f64(int64_t n, LduOp<double> A, const double* __restrict__ b,
                                                      const double* __restrict__ x, double* __restrict__ r,
                                                      double* part_out) {
    int64_t lo, hi;
    block_range(n, lo, hi);
    double acc[3] = {0.0, 0.0, 0.0};
    for (int64_t i = lo + threadIdx.x; i < hi; i += kBlock) {
        const double ri = b[i] - A.row(i, x);
        r[i] = ri;
        acc[0] += ri;
        acc[1] += ri * ri;
        acc[2] += b[i] * b[i];
    }
    block_partials<3>(acc, part_out);
}

// r -= mean (the projected residual), rounded to FP32 into r32; partials of the projected r^2
__global__ void __launch_bounds__(kBlock) k_project_round(int64_t n, double* __restrict__ r, const double* part_in,
                                                          int np_in, int project, float* __restrict__ r32,
                                                          double* part_out) {
    double tot[3];
    consume_sum<3>(part_in, np_in, tot);
    const double mean = project ? tot[0] / (double)n : 0.0;
    int64_t lo, hi;
    block_range(n, lo, hi);
    double acc[1] = {0.0};
    for (int64_t i = lo + threadIdx.x; i < hi; i += kBlock) {
        const double ri = r[i] - mean;
        r[i] = ri;
        r32[i] = (float)ri;
        acc[0] += ri * ri;
    }
    block_partials<1>(acc, part_out);
}

// out[0] = ||b||^2, out[1] = ||projected r||^2
__global__ void __launch_bounds__(kBlock) fin_mixed_norms(const double* p3, const double* p1, int nparts, State* S,
                                                          Post post) {
    double t3[3], t1[1];
    const bool lead3 = finalize_sum<3>(p3, nparts, t3);
    const bool lead1 = finalize_sum<1>(p1, nparts, t1);
    if (lead3 && lead1) {
        S->sc[S_BABS] = sqrt(t3[2]);
        S->sc[S_RNORM] = sqrt(t1[0]);
    }
    post_host(S, post);
}

// Capture mode, round 0 of the refinement: the inner tolerance tol ||b|| / ||r|| of the eager
// round (x = 0, so ||r|| is the projected ||b||), kept on the device for the inner PCG.
__global__ void k_mixed_tol_in(State* S0, double tol) {
    if (threadIdx.x != 0) return;
    const double babs = S0->sc[S_BABS], rn = S0->sc[S_RNORM];
    const double bnorm = babs != 0.0 ? babs : 1.0;
    S0->sc[S_TOLIN] = rn > 0.0 ? tol * bnorm / rn : tol;
}

// Report of a capped round 0 (S0: the norms before it, S1: after it, Sin: the inner PCG's
// state): ok when the eager refinement would have ended after round 0 with this iterate.
// Eager: rel0 < tol ends at once (the inner solve must then have been a no-op); else, with
// refinement allowed, the inner PCG runs (it must have finished within its cap, or the cap was
// maxiter), and the refinement ends after it on a breakdown at the start (x unchanged), on
// rel1 < tol, or when no further round is allowed (max_refine 1, or maxiter used up).
__global__ void k_report_mixed(const State* Sin, const State* S0, const State* S1, double tol, int ran, int full,
                               int64_t cap, int max_refine, int64_t maxiter, double* out) {
    if (threadIdx.x != 0) return;
    const double babs = S0->sc[S_BABS];
    const double bnorm = babs != 0.0 ? babs : 1.0;
    const double rel0 = S0->sc[S_RNORM] / bnorm, rel1 = S1->sc[S_RNORM] / bnorm;
    const int done = ran ? Sin->st[I_DONE] : 1;
    const int64_t iters = !ran ? 0 : done ? (int64_t)Sin->st[I_ITERS] : cap;
    const int cin = ran && done ? Sin->st[I_CONV] : 0;
    double ok = 0.0, it = (double)iters, conv = 0.0, res = rel1;
    if (rel0 < tol) {
        ok = (!ran || (done && iters == 0)) ? 1.0 : 0.0;
        it = 0.0;
        conv = 1.0;
        res = rel0;
    } else if (!ran) {
        ok = 1.0;
        res = rel0;
    } else if (done || full) {
        if (iters == 0 && !cin) {
            ok = 1.0;
            res = rel0;
        } else if (rel1 < tol) {
            ok = 1.0;
            conv = 1.0;
        } else if (max_refine == 1 || iters >= maxiter) {
            ok = 1.0;
        }
    }
    out[0] = ok;
    out[1] = it;
    out[2] = conv;
    out[3] = res;
}

// ex = (float)(d + sum_k off_k): the row sums of the FP64 operator, rounded once
__global__ void __launch_bounds__(kBlock) k_rowsum_f32(int64_t n, LduOp<double> A, float* __restrict__ ex) {
    int64_t lo, hi;
    block_range(n, lo, hi);
    for (int64_t i = lo + threadIdx.x; i < hi; i += kBlock) {
        double s = A.diag[i];
        for (int32_t k = A.ptr[i]; k < A.ptr[i + 1]; ++k) s += A.off[k];
        ex[i] = (float)s;
    }
}

static int64_t pcg_mixed_bytes(int64_t n) {
    return align256(n * 8) + align256(2 * n * 4) + align256(2 * sizeof(State)) + 2 * kPartBytes + pcg_bytes<float>(n);
}

}  // namespace

extern "C" {

int64_t ftmi_pcg_work_bytes(int64_t n) {
    const int64_t a = pcg_bytes<double>(n), b = pcg_mixed_bytes(n);
    return a > b ? a : b;
}

int ftmi_pcg_ldu(int64_t n, const double* diag, const int32_t* ptr, const int32_t* nbr, const double* off,
                 const double* b, double* x, double tol, int64_t maxiter, void* amg, int project, void* work,
                 int64_t* iters, double* res, void* stream) {
    Amg<double>* M = nullptr;
    if (amg && !(M = amg_of<double>(amg, "ftmi_pcg_ldu: the preconditioner is not an FP64 hierarchy"))) return -1;
    LduOp<double> A{diag, ptr, nbr, off};
    return pcg_run<LduOp<double>, double>(n, A, b, x, tol, maxiter, M, project, work, iters, res, (hipStream_t)stream);
}

int ftmi_pcg_ldu_sliced(int64_t n, const double* diag, const int32_t* base, const uint8_t* deg, const int32_t* nbr,
                        const double* off, const double* b, double* x, double tol, int64_t maxiter, void* amg,
                        int project, void* work, int64_t* iters, double* res, void* stream) {
    Amg<double>* M = nullptr;
    if (amg && !(M = amg_of<double>(amg, "ftmi_pcg_ldu_sliced: the preconditioner is not an FP64 hierarchy")))
        return -1;
    LduSellOp<double> A{diag, base, deg, nbr, off};
    return pcg_run<LduSellOp<double>, double>(n, A, b, x, tol, maxiter, M, project, work, iters, res,
                                              (hipStream_t)stream);
}

int ftmi_pcg_csr(int64_t n, const int32_t* indptr, const int32_t* indices, const double* data, const double* b,
                 double* x, double tol, int64_t maxiter, void* amg, int project, void* work, int64_t* iters,
                 double* res, void* stream) {
    Amg<double>* M = nullptr;
    if (amg && !(M = amg_of<double>(amg, "ftmi_pcg_csr: the preconditioner is not an FP64 hierarchy"))) return -1;
    CsrOp<double> A{indptr, indices, data};
    return pcg_run<CsrOp<double>, double>(n, A, b, x, tol, maxiter, M, project, work, iters, res, (hipStream_t)stream);
}

int ftmi_pcg_ldu_f32(int64_t n, const float* diag, const int32_t* ptr, const int32_t* nbr, const float* off,
                     const float* b, float* x, double tol, int64_t maxiter, void* amg, int project, void* work,
                     int64_t* iters, double* res, void* stream) {
    Amg<float>* M = nullptr;
    if (amg && !(M = amg_of<float>(amg, "ftmi_pcg_ldu_f32: the preconditioner is not an FP32 hierarchy"))) return -1;
    LduOp<float> A{diag, ptr, nbr, off};
    return pcg_run<LduOp<float>, float>(n, A, b, x, tol, maxiter, M, project, work, iters, res, (hipStream_t)stream);
}

int ftmi_bicgstab_ldu_mixed(int64_t n, const float* ex, const int32_t* ptr, const int32_t* nbr, const float* off,
                            const float* b, const float* inv_diag, double* x, double tol, int64_t maxiter,
                            int relative_to_initial, int b_is_residual, void* work, int64_t* iters, double* res,
                            void* stream) {
    if (n <= 0) {
        *iters = 0;
        *res = 0.0;
        return 1;
    }
    hipStream_t st = (hipStream_t)stream;
    LduDiffOp<float> A{ex, ptr, nbr, off};
    float* rb = (float*)work;
    float* d = rb + n;
    void* inner = (char*)work + align256(2 * n * (int64_t)sizeof(float));
    const float* r0 = b;  // b_is_residual: the caller formed b - A x (in higher precision)
    if (!b_is_residual) {
        k_resid_f32<<<grid_for(n), kBlock, 0, st>>>(n, A, b, x, rb);
        FTMI_LAUNCH_CHECK("ftmi_bicgstab_ldu_mixed residual");
        r0 = rb;
    }
    const int conv = bicgstab_run<LduDiffOp<float>, float>(n, A, r0, inv_diag, d, tol, maxiter, relative_to_initial,
                                                           inner, iters, res, st, true);
    if (conv < 0) return conv;
    k_add_f32<<<grid_for(n), kBlock, 0, st>>>(n, d, x);
    FTMI_LAUNCH_CHECK("ftmi_bicgstab_ldu_mixed update");
    return conv;
}

int ftmi_bicgstab_ldu_mixed_sliced(int64_t n, const float* ex, const int32_t* base, const uint8_t* deg,
                                   const int32_t* nbr, const float* off, const float* b, const float* inv_diag,
                                   double* x, double tol, int64_t maxiter, int relative_to_initial, void* work,
                                   int64_t* iters, double* res, void* stream) {
    if (n <= 0) {
        *iters = 0;
        *res = 0.0;
        return 1;
    }
    hipStream_t st = (hipStream_t)stream;
    LduSellDiffOp<float> A{ex, base, deg, nbr, off};
    float* d = (float*)work + n;
    void* inner = (char*)work + align256(2 * n * (int64_t)sizeof(float));
    const int conv = bicgstab_run<LduSellDiffOp<float>, float>(n, A, b, inv_diag, d, tol, maxiter,
                                                               relative_to_initial, inner, iters, res, st, true);
    if (conv < 0) return conv;
    k_add_f32<<<grid_for(n), kBlock, 0, st>>>(n, d, x);
    FTMI_LAUNCH_CHECK("ftmi_bicgstab_ldu_mixed_sliced update");
    return conv;
}

int ftmi_pcg_ldu_mixed(int64_t n, const double* diag, const int32_t* ptr, const int32_t* nbr, const double* off,
                       const float* off32, const double* b, double* x, double tol, int64_t maxiter, void* amg32,
                       int project, int max_refine, void* work, int64_t* iters, double* res, int64_t* rounds,
                       void* stream) {
    Amg<float>* M = nullptr;
    if (amg32 && !(M = amg_of<float>(amg32, "ftmi_pcg_ldu_mixed: the preconditioner is not an FP32 hierarchy")))
        return -1;
    Pinned* pn = pinned();
    if (!pn) {
        set_error_msg("ftmi_pcg_ldu_mixed: pinned host memory unavailable");
        return -1;
    }
    hipStream_t st = (hipStream_t)stream;
    char* w = (char*)work;
    double* r64 = (double*)w;
    w += align256(n * 8);
    float* r32 = (float*)w;
    float* ex32 = r32 + n;
    w += align256(2 * n * 4);
    State* S = (State*)w;  // S[0]: the norms of the current round; S[1]: capture mode, after round 0
    w += align256(2 * sizeof(State));
    double* pa = (double*)w;
    double* pb = (double*)(w + kPartBytes);
    void* inner = w + 2 * kPartBytes;
    LduOp<double> A64{diag, ptr, nbr, off};
    LduDiffOp<float> A32{ex32, ptr, nbr, off32};
    k_rowsum_f32<<<grid_for(n), kBlock, 0, st>>>(n, A64, ex32);
    const int g = grid_red(n);
    int64_t total = 0;
    double bnorm = 1.0, rel = 0.0;
    int conv = 0;
    *rounds = 0;
    if (g_capture) {
        // round 0 of the refinement, its inner PCG capped; the eager inner tolerance
        // tol ||b|| / ||r|| is computed on the device; the report says whether the eager solve
        // would have stopped after this round (else the host redoes the step eagerly)
        int64_t cap;
        double* report = capture_slot(maxiter, &cap);
        k_resid_f64<<<g, kBlock, 0, st>>>(n, A64, b, x, r64, pa);
        k_project_round<<<g, kBlock, 0, st>>>(n, r64, pa, g, project, r32, pb);
        fin_mixed_norms<<<1, kBlock, 0, st>>>(pa, pb, g, S, kNoPost);
        k_mixed_tol_in<<<1, 64, 0, st>>>(S, tol);
        FTMI_LAUNCH_CHECK("ftmi_pcg_ldu_mixed residual (capture)");
        const int ran = max_refine >= 1 && maxiter > 0 ? 1 : 0;
        if (ran) {
            int64_t it = 0;
            double r_in = 0.0;
            const int c = pcg_run<LduDiffOp<float>, float, double>(n, A32, r32, x, tol, maxiter, M, project, inner,
                                                                   &it, &r_in, st, false, &S->sc[S_TOLIN], cap,
                                                                   nullptr);
            if (c < 0) return c;
        }
        if (report) {
            k_resid_f64<<<g, kBlock, 0, st>>>(n, A64, b, x, r64, pa);
            k_project_round<<<g, kBlock, 0, st>>>(n, r64, pa, g, project, r32, pb);
            fin_mixed_norms<<<1, kBlock, 0, st>>>(pa, pb, g, S + 1, kNoPost);
            const State* Sin = carve_pcg<float>(n, inner).S;
            k_report_mixed<<<1, 64, 0, st>>>(Sin, S, S + 1, tol, ran, cap >= maxiter ? 1 : 0, cap, max_refine,
                                             maxiter, report);
        }
        FTMI_LAUNCH_CHECK("ftmi_pcg_ldu_mixed (capture)");
        *iters = -1;
        *res = __builtin_nan("");
        *rounds = 1;
        return 0;
    }
    for (int round = 0; round <= max_refine; ++round) {
        // the FP64 residual of the current iterate, projected, and its norm (one host read)
        k_resid_f64<<<g, kBlock, 0, st>>>(n, A64, b, x, r64, pa);
        k_project_round<<<g, kBlock, 0, st>>>(n, r64, pa, g, project, r32, pb);
        fin_mixed_norms<<<1, kBlock, 0, st>>>(pa, pb, g, S, next_post(pn, 0));
        FTMI_LAUNCH_CHECK("ftmi_pcg_ldu_mixed residual");
        if (wait_state(pn, 0, st)) return -1;
        const double babs = pn->h[0].sc[S_BABS], rnorm = pn->h[0].sc[S_RNORM];
        if (round == 0) bnorm = babs != 0.0 ? babs : 1.0;
        rel = rnorm / bnorm;
        if (rel < tol) {
            conv = 1;
            break;
        }
        if (round == max_refine || total >= maxiter) break;
        // FP32 PCG on A d = r to the absolute level tol * ||b||; its updates x += alpha p go
        // straight into the FP64 iterate (an FP32 iterate's rounding, rough from cell to cell,
        // is not cancelled by the Laplacian and would cap the FP64 residual at fine meshes)
        int64_t it = 0;
        double r_in = 0.0;
        const double tol_in = tol * bnorm / rnorm;
        const int c = pcg_run<LduDiffOp<float>, float, double>(n, A32, r32, x, tol_in, maxiter - total, M, project,
                                                               inner, &it, &r_in, st, false);
        if (c < 0) return c;
        total += it;
        ++*rounds;
        if (it == 0 && c == 0) break;  // breakdown at the start: nothing more to gain
    }
    *iters = total;
    *res = rel;
    return conv;
}

}  // extern "C"

// ------------------------------------------------------------------ allocation-free BiCGSTAB
// ftmi_bicgstab_jacobi_noalloc: the same solve as ftmi_bicgstab_jacobi (same kernels, same
// arithmetic, bit-identical x) with its workspace in a file-scope device array instead of a
// caller buffer, and no pinned status block: the host reads the device state block with a
// plain copy after every chunk of iterations. Nothing is allocated on any call. The array is
// one per device, so calls that overlap on one device race on it.
namespace {

constexpr int64_t kNoAllocRows = int64_t(1) << 18;
constexpr int64_t kNoAllocBytes = (kBiVectors * kNoAllocRows * 8 + 255) / 256 * 256 + 512 + 3 * kPartBytes;
constexpr int kNoAllocChunk = 8;  // iterations enqueued between two reads of the state block
constexpr int kNoAllocDevices = 64;

__device__ __align__(256) unsigned char scratch[kNoAllocBytes];

struct ScratchTable {
    void* ptr[kNoAllocDevices] = {};
};
static ScratchTable g_scratch;

// The device address of `scratch` on the current device, looked up once per device (the host
// name of the array is not a device address).
static void* scratch_ptr() {
    int dev = -1;
    const hipError_t e = hipGetDevice(&dev);
    if (e != hipSuccess) {
        set_error("bicgstab_noalloc: hipGetDevice", e);
        return nullptr;
    }
    if (dev < 0 || dev >= kNoAllocDevices) {
        set_error_msg("bicgstab_noalloc: device index out of range");
        return nullptr;
    }
    if (!g_scratch.ptr[dev]) {
        void* p = nullptr;
        const hipError_t r = hipGetSymbolAddress((void**)&p, HIP_SYMBOL(scratch));
        if (r != hipSuccess || !p) {
            set_error("bicgstab_noalloc: hipGetSymbolAddress(scratch)", r);
            return nullptr;
        }
        g_scratch.ptr[dev] = p;
    }
    return g_scratch.ptr[dev];
}

static int read_state(State* H, const State* S, hipStream_t st) {
    hipError_t e = hipMemcpyAsync(H, S, sizeof(State), hipMemcpyDeviceToHost, st);
    if (e == hipSuccess) e = hipStreamSynchronize(st);
    if (e != hipSuccess) {
        set_error("bicgstab_noalloc: reading the solver state", e);
        return -(int)e;
    }
    return 0;
}

template <class Op, typename T>
int bicgstab_run_noalloc(int64_t n, Op A, const T* b, const T* inv, T* x, double tol, int64_t maxiter, int rel,
                         void* work, int64_t* iters, double* res, hipStream_t st) {
    BiWork<T> W = carve_bi<T>(n, work);
    const int g = grid_red(n);
    k_state_reset<<<1, 64, 0, st>>>(W.S, W.counter);
    bi_init<Op, T, false><<<g, kBlock, 0, st>>>(n, A, b, x, W.r, W.rh, W.p[0], W.v[0], W.part[0]);
    fin_bi_init<<<1, kBlock, 0, st>>>(W.part[0], g, W.S, tol, rel, kNoPost);
    FTMI_LAUNCH_CHECK("bicgstab_noalloc init");
    State H;
    int64_t k = 1;
    for (;;) {
        // the state read here has iterations up to k - 2 finished (the last one enqueued is
        // finished by the next iteration's first kernel)
        if (int e = read_state(&H, W.S, st)) return e;
        if (H.st[I_DONE] || k > maxiter) break;
        const int64_t stop = k + kNoAllocChunk - 1 < maxiter ? k + kNoAllocChunk - 1 : maxiter;
        for (; k <= stop; ++k) bi_iteration<Op, T>(n, A, inv, x, W, g, tol, k, maxiter, kNoPost, st);
        FTMI_LAUNCH_CHECK("bicgstab_noalloc iteration");
    }
    if (k > 1) fin_bi_r<<<1, kBlock, 0, st>>>(W.part[2], g, W.S, tol, (int)(k - 1), (int)maxiter, kNoPost);
    bi_final_dev<T><<<grid_for(n), kBlock, 0, st>>>(n, x, W.x1, W.p[0], W.p[1], inv, W.S);
    FTMI_LAUNCH_CHECK("bicgstab_noalloc final");
    if (int e = read_state(&H, W.S, st)) return e;
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

int64_t ftmi_bicgstab_noalloc_max_rows(void) { return kNoAllocRows; }

int ftmi_bicgstab_jacobi_noalloc(int64_t n, const int32_t* indptr, const int32_t* indices, const double* data,
                                 const double* b, const double* inv_diag, double* x, double tol, int64_t maxiter,
                                 int relative_to_initial, int64_t* iters, double* res, void* stream) {
    if (n <= 0) {
        *iters = 0;
        *res = 0.0;
        return 1;
    }
    if (n > kNoAllocRows) {
        set_error_msg("ftmi_bicgstab_jacobi_noalloc: n exceeds ftmi_bicgstab_noalloc_max_rows()");
        return -1;
    }
    void* work = scratch_ptr();
    if (!work) return -1;
    CsrOp<double> A{indptr, indices, data};
    return bicgstab_run_noalloc<CsrOp<double>, double>(n, A, b, inv_diag, x, tol, maxiter, relative_to_initial, work,
                                                       iters, res, (hipStream_t)stream);
}

}  // extern "C"
