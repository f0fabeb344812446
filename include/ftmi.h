/* libftmi — MI355X (gfx950) finite-volume SIMPLE/SIMPLEC kernels: the C ABI.
 *
 * Conventions (changed from the reference's flowtrace/native/fvm_kernels.h:1-11 as noted):
 *   - every array argument is a DEVICE pointer (HIP global memory); every call takes a
 *     trailing `void* stream` (a hipStream_t; NULL = the default stream) and is asynchronous
 *     on it, except the Krylov solvers, which return their iteration counts and therefore
 *     wait for their own control scalars;
 *   - floating-point arrays are float64, C-contiguous, row-major (n, dim) unless a comment says
 *     otherwise; every index array is int32 (the reference uses int64 for P, N, Pb);
 *   - the caller owns and allocates every buffer (work buffers included); the library never
 *     allocates device memory;
 *   - every function returns an int status: 0 = success, < 0 = failure (text in
 *     ftmi_last_error()), except where a return value is documented (Krylov: 1 converged,
 *     0 not converged, < 0 failure);
 *   - no function is re-entrant on a shared work buffer; one stream per solver.
 *
 * Face adjacency used by the fused SIMPLEC stages: for cell c, entries cptr[c]..cptr[c+1]-1
 * list its interior faces in increasing face order, cface[k] = (f << 1) | (c is f's
 * neighbour), cnbr[k] = the cell across face f; bptr/bface list its boundary faces.
 * A cell-cell matrix in this layout ("LDU" form) is diag[nc] plus off[k], the coefficient
 * of cnbr[k] in row c.
 */
#ifndef FTMI_H
#define FTMI_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define FTMI_ABI_VERSION 107

/* replaces ft_abi_version (flowtrace/native/fvm_kernels.h:27) */
int ftmi_abi_version(void);
/* new: text of the last failure on the calling thread */
const char* ftmi_last_error(void);
/* replaces ft_max_threads (fvm_kernels.h:28): HIP devices visible */
int ftmi_device_count(void);

/* ---------------------------------------------------------------- drop-in operators */

/* replaces ft_csr_spmm (fvm_kernels.h:31): Y[nrows, k] = A @ X[ncols, k] */
int ftmi_csr_spmm(int64_t nrows, int64_t k, const int32_t* indptr, const int32_t* indices, const double* data,
                  const double* X, double* Y, void* stream);

/* replaces ft_face_interp (fvm_kernels.h:35): out[f, :] = w[f] phi[P[f], :] + (1 - w[f]) phi[N[f], :] */
int ftmi_face_interp(int64_t nf, int64_t k, const double* w, const int32_t* P, const int32_t* N, const double* phi,
                     double* out, void* stream);

/* replaces ft_ls_gradient (fvm_kernels.h:39): least-squares gradient over the Cdiff (iptr, iidx)
 * and Cb (bptr, bidx) adjacencies, out[c, :dim] = Ginv[c] r[c] */
int ftmi_ls_gradient(int64_t nc, int64_t dim, const int32_t* iptr, const int32_t* iidx, const int32_t* bptr,
                     const int32_t* bidx, const int32_t* P, const int32_t* N, const double* wl, const double* d,
                     const double* wlb, const double* db, const double* Ginv, const double* phi, const double* phib,
                     double* out, void* stream);

/* replaces ft_gg_gradient (fvm_kernels.h:48): Green-Gauss gradient with optional skewness
 * correction grad_f . dxm_f (grad may be NULL) */
int ftmi_gg_gradient(int64_t nc, int64_t dim, const int32_t* iptr, const int32_t* iidx, const double* isgn,
                     const int32_t* bptr, const int32_t* bidx, const int32_t* P, const int32_t* N, const double* w,
                     const double* S, const double* Sb, const double* vol, const double* dxm, const double* phi,
                     const double* phib, const double* grad, double* out, void* stream);

/* replaces ft_assemble (fvm_kernels.h:56): CSR data from diag and the face coefficients.
 * The reference scatters through slot_PN/slot_NP in one sequential face loop; here each row
 * is filled by one thread from a per-row list built once by ftmi_assemble_plan_host from the
 * same slots: src[src_ptr[c]..src_ptr[c+1]) holds pairs (face << 1 | which, slot), which 0 =
 * aPN, 1 = aNP, in face order. */
int ftmi_assemble(int64_t nnz, int64_t nc, int64_t nf, const int32_t* indptr, const int32_t* slot_diag,
                  const int32_t* src_ptr, const int32_t* src, const double* diag, const double* aPN,
                  const double* aNP, double* data, void* stream);
/* host helper (host pointers): builds src_ptr[nc+1] and src[4 nf] from the slots */
int ftmi_assemble_plan_host(int64_t nc, int64_t nf, const int32_t* indptr, const int32_t* slot_PN,
                            const int32_t* slot_NP, int32_t* src_ptr, int32_t* src);

/* replaces ft_bicgstab_jacobi (fvm_kernels.h:64): Jacobi-right-preconditioned BiCGSTAB with the
 * stopping rules of flowtrace/linalg/krylov.py:68-121. x: x0 in, solution out. work: device
 * buffer of ftmi_bicgstab_work_bytes(n) bytes (the reference: 7 n doubles). *iters, *res are
 * host pointers. Returns 1 converged, 0 not, < 0 error; *res = NaN on breakdown. */
int64_t ftmi_bicgstab_work_bytes(int64_t n);
int ftmi_bicgstab_jacobi(int64_t n, const int32_t* indptr, const int32_t* indices, const double* data,
                         const double* b, const double* inv_diag, double* x, double tol, int64_t maxiter,
                         int relative_to_initial, void* work, int64_t* iters, double* res, void* stream);

/* new, no counterpart in the reference: ftmi_bicgstab_jacobi for callers that cannot allocate
 * (code under HIP graph capture, tight serving loops, harnesses that forbid hipMalloc). Same
 * solve and the same x, bit for bit, without the work argument: the workspace is a file-scope
 * device array of the library (about 23 MB per device, held until the process ends), and the
 * host reads the solver state with a plain device-to-host copy every 8 iterations. No call
 * allocates device or host memory. n may not exceed ftmi_bicgstab_noalloc_max_rows() (262144;
 * larger n returns -1 with a message). All calls on one device share the array: calls that
 * overlap on one device (two streams, two threads) race and must be serialised by the caller. */
int64_t ftmi_bicgstab_noalloc_max_rows(void);
int ftmi_bicgstab_jacobi_noalloc(int64_t n, const int32_t* indptr, const int32_t* indices, const double* data,
                                 const double* b, const double* inv_diag, double* x, double tol, int64_t maxiter,
                                 int relative_to_initial, int64_t* iters, double* res, void* stream);

/* new: opt-in timing of the BiCGSTAB p-update + SpMV kernel (HIP events around each launch on
 * its stream); enable resets the totals, read waits for the recorded events */
/* new: capture mode (1) for recording solves in a HIP graph: the Krylov solvers enqueue every
 * iteration up to maxiter (or the plan's cap) without reading the device state back (a
 * finished solve's kernels return at once on the device), report *iters = -1 and *res = NaN,
 * and the mixed pressure solve applies exactly one FP32 correction. Returns the previous mode. */
int ftmi_capture_mode(int on);
/* new: capture plan. The solves enqueued in capture mode take slots 0, 1, ... in call order
 * (a mixed pressure solve takes one); slot s enqueues min(maxiter, caps[s]) iterations
 * (caps[s] <= 0, or s >= ncaps: maxiter) and, when report != NULL and s < slots, writes
 * four doubles to the device buffer report + 4 s when it ends: ok (1 when the enqueued
 * iterations gave the eager solve's result), iterations, converged (0/1) and the residual
 * the eager solve reports. report = NULL clears the plan. Resets the slot counter. */
int ftmi_capture_plan(const int64_t* caps, int ncaps, double* report, int slots);
/* new: slots taken since the last ftmi_capture_plan */
int ftmi_capture_slots_used(void);
/* new: workgroup cap of the streaming kernels (default 2048 = 8 per CU; 64..65536), the knob
 * the autotuner sets; returns the previous cap */
int ftmi_set_grid_cap(int cap);
int ftmi_timer_enable(int on);
int ftmi_timer_read(double* total_ms, int64_t* launches);

/* ---------------------------------------------------------------- Krylov on the LDU form */

/* new (the reference's momentum/energy solve, simple.py:489-497, on the face layout) */
/* new: BiCGSTAB on the face operator with its half-edges in 64-row slices (k-th half-edge of row
 * i at base[i / 64] + 64 k + i % 64, k < deg[i]); same algorithm and sums as ftmi_bicgstab_ldu */
int ftmi_bicgstab_ldu_sliced(int64_t n, const double* diag, const int32_t* base, const uint8_t* deg,
                             const int32_t* nbr, const double* off, const double* b, const double* inv_diag, double* x,
                             double tol, int64_t maxiter, int relative_to_initial, void* work, int64_t* iters,
                             double* res, void* stream);
/* new: soff[pos[k]] = off[k], half-edge values into the slice layout */
int ftmi_ldu_slice(int64_t nhe, const int32_t* pos, const double* off, double* soff, void* stream);
int ftmi_ldu_slice_f32(int64_t nhe, const int32_t* pos, const float* off, float* soff, void* stream);
int ftmi_bicgstab_ldu(int64_t n, const double* diag, const int32_t* ptr, const int32_t* nbr, const double* off,
                      const double* b, const double* inv_diag, double* x, double tol, int64_t maxiter,
                      int relative_to_initial, void* work, int64_t* iters, double* res, void* stream);

/* new: PCG of flowtrace/linalg/krylov.py:25-65 from x0 = 0 (x is output only), preconditioned
 * by one V-cycle of the hierarchy `amg` (NULL: none); project != 0 removes the mean of the
 * residual and of every preconditioned residual (the singular pressure system,
 * simple.py:629-634). work: ftmi_pcg_work_bytes(n) bytes. */
int64_t ftmi_pcg_work_bytes(int64_t n);
int ftmi_pcg_ldu(int64_t n, const double* diag, const int32_t* ptr, const int32_t* nbr, const double* off,
                 const double* b, double* x, double tol, int64_t maxiter, void* amg, int project, void* work,
                 int64_t* iters, double* res, void* stream);
/* new: ftmi_pcg_ldu with the operator's half-edges in the slice layout of
 * ftmi_bicgstab_ldu_sliced (the same algorithm and sums, coalesced loads) */
int ftmi_pcg_ldu_sliced(int64_t n, const double* diag, const int32_t* base, const uint8_t* deg, const int32_t* nbr,
                        const double* off, const double* b, double* x, double tol, int64_t maxiter, void* amg,
                        int project, void* work, int64_t* iters, double* res, void* stream);
int ftmi_pcg_csr(int64_t n, const int32_t* indptr, const int32_t* indices, const double* data, const double* b,
                 double* x, double tol, int64_t maxiter, void* amg, int project, void* work, int64_t* iters,
                 double* res, void* stream);

/* ---------------------------------------------------------------- mixed precision (new; the
 * reference runs one precision throughout, simple.py:77). FP32 matrices, vectors and
 * arithmetic inside the Krylov iterations; FP64 where the outer iteration needs it. */
/* FP32 PCG with an FP32 hierarchy (ftmi_amg_create_f32); work: ftmi_pcg_work_bytes(n) */
int ftmi_pcg_ldu_f32(int64_t n, const float* diag, const int32_t* ptr, const int32_t* nbr, const float* off,
                     const float* b, float* x, double tol, int64_t maxiter, void* amg, int project, void* work,
                     int64_t* iters, double* res, void* stream);
/* transport solve in correction form: rb = b - A x in FP32 (or, with b_is_residual, rb = b as
 * given: the residual the FP32 transport assembly forms in difference form), Jacobi-BiCGSTAB in
 * FP32 on A d = rb from d = 0 (the same iterates and stopping rule as the solve from x, with the
 * absolute floor 1e-14 ||rb||), then x += d in FP64. A is applied in difference form,
 * ex_i x_i + sum_k off_k (x_nbr - x_i), with ex the row sums (ftmi_assemble_args.ex32).
 * work: ftmi_bicgstab_work_bytes(n) */
int ftmi_bicgstab_ldu_mixed(int64_t n, const float* ex, const int32_t* ptr, const int32_t* nbr, const float* off,
                            const float* b, const float* inv_diag, double* x, double tol, int64_t maxiter,
                            int relative_to_initial, int b_is_residual, void* work, int64_t* iters, double* res,
                            void* stream);
/* ftmi_bicgstab_ldu_mixed with b_is_residual, the operator's half-edges in the slice layout of
 * ftmi_bicgstab_ldu_sliced (the same sums, coalesced loads) */
int ftmi_bicgstab_ldu_mixed_sliced(int64_t n, const float* ex, const int32_t* base, const uint8_t* deg,
                                   const int32_t* nbr, const float* off, const float* b, const float* inv_diag,
                                   double* x, double tol, int64_t maxiter, int relative_to_initial, void* work,
                                   int64_t* iters, double* res, void* stream);
/* pressure solve by FP64 iterative refinement: r = b - A x in FP64 (projected when
 * `project`), stop when ||r|| / ||b|| < tol, else FP32 AMG-PCG on A32 d = r to the level
 * tol ||b||, its updates added straight into the FP64 x, at most max_refine times (FP32
 * matrix, hierarchy and Krylov vectors; FP64 iterate). A32 is applied in difference form,
 * ex_i d_i + sum_k off32_k (d_nbr - d_i), with ex the FP64 row sums rounded once, so FP32
 * rounding does not swamp the residual of the nearly singular pressure Laplacian.
 * off32 = (float) off. *iters: FP32 PCG iterations in total; *res: the final FP64 relative
 * residual; *rounds: corrections applied. Returns 1 if the FP64 residual met tol.
 * work: ftmi_pcg_work_bytes(n) */
int ftmi_pcg_ldu_mixed(int64_t n, const double* diag, const int32_t* ptr, const int32_t* nbr, const double* off,
                       const float* off32, const double* b, double* x, double tol, int64_t maxiter, void* amg32,
                       int project, int max_refine, void* work, int64_t* iters, double* res, int64_t* rounds,
                       void* stream);
/* y = (float) x */
int ftmi_cast_f64_f32(int64_t n, const double* x, float* y, void* stream);
/* new: y[j] = x[idx[j]] (values of a transposed pattern in the AMG numeric refresh) */
int ftmi_gather_f64(int64_t n, const int64_t* idx, const double* x, double* y, void* stream);

/* ---------------------------------------------------------------- smoothed-aggregation AMG
 * replaces SmoothedAggregationAMG (flowtrace/linalg/amg.py:80-146). The hierarchy handle
 * holds device pointers owned by the caller; levels are added finest first; the coarsest
 * system is solved with a dense inverse (row-major). Level vectors b, x, y, r are caller-
 * owned device buffers of the level's size. */
void* ftmi_amg_create(double omega, int sweeps);
void ftmi_amg_destroy(void* amg);
/* new: shift (-3..3) of the V-cycle's threads-per-row choice (1, 4, 16 or 64 lanes per row
 * from the level's mean row length), applied to levels added afterwards and by
 * ftmi_amg_retune; 0 (the default) keeps the row-length rule. Returns the previous shift. */
int ftmi_set_tpr_bias(int bias);
/* new: re-pick the threads per row of every level of hierarchy h with the current shift */
int ftmi_amg_retune(void* h);
/* nnz_a, nnz_p: entries of A and P (the last row pointers), given so that adding a level
 * reads nothing back from the device */
int ftmi_amg_add_level(void* amg, int64_t n, int64_t nnz_a, int64_t nnz_p, const int32_t* a_ptr, const int32_t* a_col,
                       const double* a_val,
                       const int32_t* p_ptr, const int32_t* p_col, const double* p_val, const int32_t* r_ptr,
                       const int32_t* r_col, const double* r_val, const double* dinv, double* vb, double* vx,
                       double* vy, double* vr);
int ftmi_amg_set_coarse(void* amg, int64_t n, const double* inv, double* vb, double* vx);
/* new: instead of a dense inverse, the coarsest system (CSR a_*, nnz entries, inverse diagonal
 * dinv) is solved by `sweeps` >= 2 Jacobi sweeps from zero with weight omega (below
 * 2 / rho(D^-1 A) for them to converge), for a hierarchy cut at a level count whose coarsest
 * matrix is too large for a dense inverse; vb, vx, vy: its vectors */
int ftmi_amg_set_coarse_iter(void* amg, int64_t n, int64_t nnz, const int32_t* a_ptr, const int32_t* a_col,
                             const double* a_val, const double* dinv, int sweeps, double omega, double* vb,
                             double* vx, double* vy);
int ftmi_amg_vcycle(void* amg, const double* b, double* x, void* stream);
/* the same hierarchy with FP32 values and vectors (new: the mixed-precision pressure solve);
 * a handle takes only the entry points of its own precision */
void* ftmi_amg_create_f32(double omega, int sweeps);
int ftmi_amg_add_level_f32(void* amg, int64_t n, int64_t nnz_a, int64_t nnz_p, const int32_t* a_ptr,
                           const int32_t* a_col, const float* a_val,
                           const int32_t* p_ptr, const int32_t* p_col, const float* p_val, const int32_t* r_ptr,
                           const int32_t* r_col, const float* r_val, const float* dinv, float* vb, float* vx,
                           float* vy, float* vr);
int ftmi_amg_set_coarse_f32(void* amg, int64_t n, const float* inv, float* vb, float* vx);
int ftmi_amg_set_coarse_iter_f32(void* amg, int64_t n, int64_t nnz, const int32_t* a_ptr, const int32_t* a_col,
                                 const float* a_val, const float* dinv, int sweeps, double omega, float* vb, float* vx,
                                 float* vy);
int ftmi_amg_vcycle_f32(void* amg, const float* b, float* x, void* stream);
/* host helper (host pointers): three-pass greedy aggregation of amg.py:26-57 over a strength
 * graph in CSR form; returns the number of aggregates */
int64_t ftmi_amg_aggregate_host(int64_t n, const int64_t* ptr, const int64_t* idx, int64_t* agg);
/* host helper: the same aggregation over a CSR pattern with a strength mask per entry
 * (int32 in/out); size[a] = members of aggregate a (size holds n entries) */
int64_t ftmi_amg_aggregate_mask_host(int64_t n, const int32_t* ptr, const int32_t* col, const uint8_t* mask,
                                     int32_t* agg, int32_t* size);
/* new: the same aggregation over the strength graph as its own CSR (s_ptr, s_col: the strong
 * entries of each row in column order, from ftmi_amgs_strong_count / _fill) */
int64_t ftmi_amg_aggregate_csr_host(int64_t n, const int32_t* s_ptr, const int32_t* s_col, int32_t* agg,
                                    int32_t* size);

/* new: device reductions of the outer-iteration diagnostics (flowtrace/unstructured/diagnostics.py:
 * 103-181 _rows_of_system / _row_set, 214-290 evaluate); results land in device doubles.
 * rows: r = b' - A x, s = |b'| + |A||x| of an LDU system, b' = b - c dgb xb when dgb != NULL */
int ftmi_diag_rows(int64_t n, const double* dg, const int32_t* ptr, const int32_t* nbr, const double* off,
                   const double* b, const double* dgb, double c, const double* xb, const double* x, double* r,
                   double* s, void* stream);
/* per cell: m = net outflow of (F, Fb), t = sum |flux| (cface = f << 1 | is-neighbour) */
int ftmi_diag_cell_flux(int64_t n, const int32_t* cptr, const int32_t* cface, const int32_t* bptr,
                        const int32_t* bface, const double* F, const double* Fb, double* m, double* t, void* stream);
/* doubles of work for the two reductions below */
int64_t ftmi_diag_work_doubles(void);
/* a row set of ncomp (<= 4) (r, s) vectors: out[0] = s_ref, out[1 + 9 c + j] = row-scaled max,
 * its row (-1 if none), active, excluded, excluded max |r|, non-finite, l1, max, l2 */
int ftmi_diag_rowset(int64_t n, int ncomp, const double* const* r, const double* const* s, double gamma_k,
                     double* out, double* work, void* stream);
/* a vector field (ncomp components stride apart): out[0..7] = max row norm of x - y (y may be
 * NULL), max row norm of x, min, max, non-finite entries, sum, sum |x|, sum x^2 */
int ftmi_diag_vstats(int64_t n, int ncomp, int64_t stride, const double* x, const double* y, double* out,
                     double* work, void* stream);

/* new: bulk number parsing for the ASCII sections of a Fluent mesh file (host only; the
 * reference parses them in Python, flowtrace/mesh/fluent_msh.py:151-156 and 260). Blank-separated
 * tokens; returns the count, -1 - k if token k is malformed, -2 - cap if there are more than cap. */
int64_t ftmi_parse_hex(const char* text, int64_t len, int64_t* out, int64_t cap);
int64_t ftmi_parse_f64(const char* text, int64_t len, double* out, int64_t cap);
/* new: the top-level section table of a Fluent mesh file (the section grammar of
 * flowtrace/mesh/fluent_msh.py:196-350, read natively): per section 6 int64 (index, header
 * begin, header end, body begin, body end, flags; bit 0 = has a data group). Returns the section
 * count, -100 - count when cap is too small, or a negative error code (1 expected parenthesis,
 * 2 nested group in grid data, 3 unterminated group, 4 unterminated string, 5 no end marker of a
 * binary section) with the byte position in *err_pos. */
int64_t ftmi_fluent_index(const char* buf, int64_t len, int64_t* table, int64_t cap, int64_t* err_pos);
/* new: record starts of a mixed face section (each record: k, k nodes, c0, c1); returns the
 * integers used or -1 when the records run past nvals */
int64_t ftmi_fluent_walk(const int64_t* vals, int64_t nvals, int64_t nrec, int64_t* starts);

/* new: SA-AMG set-up on the device (amg.py:60-117, everything but the aggregation) */
/* CSR data of the pressure matrix from its face form (he_slot: CSR slot of each half-edge) */
int ftmi_amgs_ldu_to_csr(int64_t n, const int32_t* cptr, const int32_t* he_slot, const int32_t* slot_diag,
                         const int32_t* indptr, const double* diag, const double* off, double* data, void* stream);
/* diagonal and its inverse */
int ftmi_amgs_diag(int64_t n, int64_t nnz, const int32_t* ptr, const int32_t* col, const double* val, double* diag,
                   double* dinv, void* stream);
/* strength mask |a_ij| >= theta sqrt(|a_ii a_jj|), i != j (amg.py:60-66) */
int ftmi_amgs_strength(int64_t n, int64_t nnz, const int32_t* ptr, const int32_t* col, const double* val,
                       const double* diag, double theta, uint8_t* mask, void* stream);
/* new: the strength mask as above, compared entry by entry with a previous mask old of the
 * same pattern; where any entry differs, gen is stored into *flag (device memory). *flag is
 * never cleared: pass a gen it does not hold (a counter), then gen == *flag means "differs".
 * Lets an AMG rebuild find the levels whose aggregation would change without a host pass. */
int ftmi_amgs_strength_cmp(int64_t n, int64_t nnz, const int32_t* ptr, const int32_t* col, const double* val,
                           const double* diag, double theta, uint8_t* mask, const uint8_t* old, int32_t* flag,
                           int gen, void* stream);
/* new: *flag = gen where the n 64-bit words at a and b differ anywhere (bitwise); *flag is never
 * cleared (pass a gen it does not hold). Tells whether a pressure matrix is exactly the one an
 * AMG hierarchy was built for, so that the hierarchy can be reused as it is. */
int ftmi_differs_u64(int64_t n, const void* a, const void* b, int32_t* flag, int gen, void* stream);
/* new: the strength graph of a level as its own CSR: per-row counts of the mask's strong
 * entries, then (after ftmi_scan_i32 of the counts into s_ptr) their columns in order */
int ftmi_amgs_strong_count(int64_t n, const int32_t* ptr, const uint8_t* mask, int32_t* cnt, void* stream);
int ftmi_amgs_strong_fill(int64_t n, const int32_t* ptr, const int32_t* col, const uint8_t* mask,
                          const int32_t* s_ptr, int32_t* s_col, void* stream);
/* tentative prolongator as CSR: t_ptr = 0..n, t_val = 1/sqrt(|aggregate|) */
int ftmi_amgs_tentative(int64_t n, const int32_t* agg, const int32_t* agg_size, int32_t* t_ptr, double* t_val,
                        void* stream);
/* rho(D^-1 A) by power iteration from x (amg.py:69-77); y: 2 n doubles of work; partials:
 * 4096 doubles; out[0] = the estimate (device) */
int ftmi_amgs_power(int64_t n, int64_t nnz, const int32_t* ptr, const int32_t* col, const double* val,
                    const double* dinv, double* x, double* y, int iters, double* partials, double* out, void* stream);
/* deterministic sparse product C = diag(xscale) X Y (nrows x ncols), count (fill = 0) or fill
 * pass; with agg: C = T - (omega_scale / *omega_num) diag(xscale) X Y (the smoothed
 * prolongator). Columns come out sorted and each entry is summed in the order of X's row.
 * long_rows (nrows int32) and n_long (1 int32, zero before the count pass) are work shared by
 * the two passes: the count pass lists the rows of X too long for one thread there, and both
 * passes read that list. */
int ftmi_amgs_spgemm(int fill, int64_t nrows, int64_t ncols, const int32_t* x_ptr, const int32_t* x_col,
                     const double* x_val,
                     const int32_t* y_ptr, const int32_t* y_col, const double* y_val, const double* xscale,
                     const int32_t* agg, const double* tval, const double* omega_num, double omega_scale,
                     int32_t* rowcnt, const int32_t* c_ptr, int32_t* c_col, double* c_val, int* overflow,
                     int32_t* long_rows, int32_t* n_long, void* stream);
/* exclusive scan: out[0..n] (out[n] = total); bsum: 2048 int64 of work */
int ftmi_scan_i32(int64_t n, const int32_t* in, int32_t* out, int64_t* bsum, void* stream);
/* transpose: column counts, then placement and per-row sort */
int ftmi_amgs_count_cols(int64_t nnz, const int32_t* col, int32_t* cnt, void* stream);
int ftmi_amgs_transpose_fill(int64_t nrows, int64_t ncols, const int32_t* ptr, const int32_t* col, const double* val,
                             const int32_t* r_ptr, int32_t* fill, int32_t* r_col, double* r_val, void* stream);

/* ---------------------------------------------------------------- fused SIMPLEC stages */

typedef struct ftmi_mesh {
    int64_t nc, nf, nb, dim;
    /* interior faces */
    const int32_t *P, *N;
    const double *w, *alpha, *wl;                       /* (nf) */
    const double *S, *d, *Tvec, *dP_f, *dN_f, *dxm;     /* (nf, dim) */
    /* boundary faces */
    const int32_t* Pb;
    const double *Sb, *db, *db_t, *nb_unit;             /* (nb, dim) */
    const double *alpha_b, *wlb, *Sb_mag;               /* (nb) */
    /* cells */
    const double *vol, *Ginv;                           /* (nc), (nc, dim, dim) */
    /* adjacency (see the file comment) */
    const int32_t *cptr, *cface, *cnbr, *bptr, *bface;
} ftmi_mesh;

typedef struct ftmi_bc {
    const int32_t* kind;      /* (nb): 0 wall, 1 inlet, 2 outlet, 3 pressure, 4 symmetry */
    const int32_t* tkind;     /* (nb): 0 adiabatic, 1 fixed T, 2 heat flux */
    const double* ub_fixed;   /* (nb, dim) */
    const double* pb_fixed;   /* (nb) */
    const double* tval;       /* (nb) */
} ftmi_bc;

/* reduction scratch: partials of >= 2048 * 8 doubles and a zero-initialised counter */
typedef struct ftmi_red {
    double* partials;
    unsigned* counter;
} ftmi_red;

typedef struct ftmi_assemble_args {
    const double *F, *Fb;
    double gamma;
    int energy;               /* 0 momentum (dim fields), 1 energy (one field) */
    int nfields;
    const double* phi[3];     /* (nc) each */
    const double* grad[3];    /* (nc, dim) each */
    const double* phib[3];    /* (nb) each */
    const double* gp;         /* momentum: pressure gradient (nc, dim) */
    double drive[3];
    int has_drive;
    double relax;
    int second_order;
    int simplec;              /* 1 SIMPLEC, 0 SIMPLE, -1 no dP */
    double cp;
    double *diag, *diag_r, *off, *b[3], *dP;
    double* resid;            /* device, 2 nfields doubles */
    double* inv_r;            /* 1 / diag_r (may be NULL) */
    double* x0[3];            /* copies of phi[q], the Krylov start vectors (may be NULL) */
    int out_fp32;             /* 1: FP32 arithmetic, and diag/diag_r/off/b/inv_r go to the FP32
                                 arrays below (the mixed-precision transport systems; b32 then
                                 holds the residual b_r - A_r phi of the relaxed system at phi,
                                 formed in difference form); dP, resid and x0 stay FP64 */
    float *diag32, *diag_r32, *off32, *b32[3], *inv_r32;
    float* ex32;              /* out_fp32: row sums diag_r + sum off of the relaxed operator,
                                 built from the face fluxes (the difference form) */
    /* face-varying diffusivity (simple.py:333-338, the (interior, boundary) form of gamma):
     * gam_f (nf) and gam_b (nb), or NULL for `gamma` everywhere */
    const double *gam_f, *gam_b;
    const double* extra[3];   /* momentum: a further per-cell source of field q, added after
                                 the body force (the volumetric source of simple.py:535-536, or
                                 the turbulent transpose stress of simple.py:546-548); energy:
                                 extra[0] is added to the heat input after the flux walls
                                 (simple.py:795-796); NULL: none */
    double* cell_resid;       /* per-cell residual terms, or NULL: (2 nfields, nc) doubles,
                                 row 2q |rhs - A phi| and row 2q+1 |diag phi| of field q, whose
                                 sums over a case are that case's residual sums (batches) */
} ftmi_assemble_args;


/* new: out[r nseg + k] = sum (or sum of |x| when absval) of x[r stride + i] over i in
 * [offsets[k], offsets[k+1]), r < nrows (offsets: device, nseg + 1 entries, nondecreasing,
 * offsets[nseg] <= stride); deterministic. The per-case sums of a batch on one mesh. */
int ftmi_segment_sums(int64_t nseg, const int64_t* offsets, int64_t nrows, int64_t stride, const double* x,
                      int absval, double* out, void* stream);

/* ---------------------------------------------------------------- k-omega SST
 * (flowtrace/unstructured/turbulence.py:116-437; new entry points, the reference computes the
 * model in NumPy/CuPy array code). Products of constants are formed by the caller in the
 * reference's left-to-right order so that the device arithmetic is the reference's. */
typedef struct ftmi_sst_coef {
    double rho, mu;
    double two_rho_sw2;    /* 2 rho sigma_w2 */
    double four_rho_sw2;   /* 4 rho sigma_w2 */
    double nu500;          /* 500 nu */
    double ten_bstar_rho;  /* 10 beta* rho */
    double bstar_rho;      /* beta* rho */
    double rho_a1;         /* rho a1 */
    double g1, g2;         /* gamma_1, gamma_2 */
} ftmi_sst_coef;

/* k-omega SST: boundary values of k and omega, fixed (1) at inlets and backflow faces
 * (turbulence.py:263-269) */
int ftmi_sst_bvals(const ftmi_mesh* mesh, const ftmi_bc* bc, const double* Fb, const double* k, const double* w,
                   const double* k_in, const double* w_in, double* kb, double* wb, double* fixed, void* stream);
/* strain rate S, and per cell the diffusivities and the volume-weighted sources of the k and
 * omega transports from gU (dim, nc, dim) and the LS gradients gk, gw (nc, dim)
 * (turbulence.py:99-113, 367-380, 419-433) */
int ftmi_sst_cell(const ftmi_mesh* mesh, const ftmi_sst_coef* coef, const double* gU, const double* gk,
                  const double* gw, const double* k, const double* w, const double* mut, const double* y,
                  double* S, double* gam_k, double* sck, double* spk, double* gam_w, double* scw, double* spw,
                  void* stream);
/* doubles of `work` that ftmi_sst_assemble needs */
int ftmi_sst_work_doubles(void);
/* one SST transport system (turbulence.py:271-304): LDU coefficients (off in the cell-centric
 * half-edge order), diag, rhs, the relaxed diag_r / b / inv_r, x0 = phi; out3 (device): sum
 * |rhs - A phi| and sum |diag phi| over the cells not at the floor and not `held`, max diag */
int ftmi_sst_assemble(const ftmi_mesh* mesh, const double* F, const double* Fb, const double* gam, const double* phi,
                      const double* grad, const double* phib, const double* fixed, const double* Sc, const double* Sp,
                      const uint8_t* held, double floor_v, double relax, double* diag, double* rhs, double* diag_r,
                      double* off, double* b, double* inv_r, double* x0, double* work, double* out3, void* stream);
/* the wall cells' omega held by the penalty 1e10 max(diag) (turbulence.py:286-290, 390-397);
 * k is the field before the update, out3 the assembly's */
int ftmi_sst_hold(int64_t nh, const int32_t* cells, const double* y_wall_cell, const double* k, const double* phi,
                  const double* out3, double nu, double relax, const double* diag, const double* rhs, double* diag_r,
                  double* b, double* inv_r, void* stream);
/* floors, then mu_t = rho a1 k / max(a1 omega, S F2) of the new fields (turbulence.py:382-388) */
int ftmi_sst_update(int64_t n, const double* xk, const double* xw, double kmin, double wmin, const double* y,
                    const double* S, const ftmi_sst_coef* coef, double* k, double* w, double* mut, void* stream);
/* face diffusivities: mode 0 momentum (c0 = mu; mu + mu_t), mode 1 energy (c0 = k/cp; +
 * mu_t/Pr_t, none on walls) (turbulence.py:235-249) */
int ftmi_sst_face_gamma(const ftmi_mesh* mesh, const ftmi_bc* bc, const double* mut, double c0, int mode,
                        double* gam_f, double* gam_b, void* stream);
/* wall-function viscosity on wall faces of gam_b and y+ (nb) from Spalding's law
 * (turbulence.py:65-87, 399-417); iters: one device int of work */
int ftmi_sst_wall(const ftmi_mesh* mesh, const ftmi_bc* bc, const double* U, const double* yb, double mu, double nu,
                  int* iters, double* gam_b, double* yplus, void* stream);
/* explicit div(mu_t (grad U)^T) per cell, out (dim, nc) (turbulence.py:251-261) */
int ftmi_sst_tstress(const ftmi_mesh* mesh, const double* mut, const double* gU, double* out, void* stream);

/* new: boundary velocities and pressures (simple.py:305-325); U is (dim, nc) component-major,
 * gU is (dim, nc, dim), ub is (dim, nb); gU/gp may be NULL (no lagged reconstruction) */
int ftmi_simplec_bvals(const ftmi_mesh* mesh, const ftmi_bc* bc, const double* U, const double* p, const double* gU,
                       const double* gp, double* ub, double* pb, void* stream);
/* new: boundary temperatures (simple.py:783-791) */
int ftmi_energy_bvals(const ftmi_mesh* mesh, const ftmi_bc* bc, const double* T, const double* gT, const double* Fb,
                      double* Tb, void* stream);
/* new: up to four least-squares gradients in one pass (fvmesh.py:268-286 / fvm_kernels.cpp:81-105);
 * phi/phib/out are host arrays of device pointers; phib == NULL or phib[q] == NULL means the
 * pressure-correction boundary value (1 - m_pres) phi_P with m_pres from bkind */
int ftmi_ls_gradients(const ftmi_mesh* mesh, int nfields, const double* const* phi, const double* const* phib,
                      double* const* out, const int32_t* bkind, void* stream);
/* new: Green-Gauss gradient on the face adjacency (fvmesh.py:288-309) */
int ftmi_gg_gradient_fused(const ftmi_mesh* mesh, const double* phi, const double* phib, const double* grad,
                           double* out, void* stream);
/* new: coefficients + right-hand sides + residual sums + relaxed system + SIMPLEC dP
 * (simple.py:340-387, 526-590, 778-809) */
int ftmi_transport_assemble(const ftmi_mesh* mesh, const ftmi_bc* bc, const ftmi_assemble_args* args,
                            const ftmi_red* red, void* stream);
/* new: Rhie-Chow face fluxes, pressure-correction system, mass imbalance (simple.py:590-621);
 * cont_out: device scalar sum |mass| */
int ftmi_rhie_chow(const ftmi_mesh* mesh, const ftmi_bc* bc, double rho, const double* Ustar, const double* p,
                   const double* gp, const double* dP, double* Fst, double* cf, double* Fbst, double* cb,
                   double* diag_p, double* offp, double* mass, double* neg_mass, double* cont_out,
                   const ftmi_red* red, void* stream);
/* new: flux, pressure and velocity corrections (simple.py:639-646); sums_out (device, 3):
 * sum p_new (before mean removal), sum |F|, sum |Fb| */
int ftmi_correct(const ftmi_mesh* mesh, const ftmi_bc* bc, const double* pp, const double* Ustar, const double* dP,
                 const double* Fst, const double* cf, const double* Fbst, const double* cb, double relax_p,
                 int remove_mean, double* U, double* p, double* F, double* Fb, double* gpp_out, double* sums_out,
                 const ftmi_red* red, void* stream);
/* new: initial face fluxes from the cell velocity (simple.py:110-112) */
int ftmi_init_flux(const ftmi_mesh* mesh, const ftmi_bc* bc, double rho, const double* U, const double* ub,
                   double* F, double* Fb, void* stream);
/* new: *out = sum |x| (device scalar) */
int ftmi_abs_sum(int64_t n, const double* x, double* out, const ftmi_red* red, void* stream);

#ifdef __cplusplus
}
#endif

#endif
