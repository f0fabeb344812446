// Fused stage kernels of one SIMPLE/SIMPLEC outer iteration (flowtrace/unstructured/simple.py:500-703).
//
// Every face loop runs cell-centric: a cell walks its interior faces through the adjacency
// (cptr, cface, cnbr), faces in increasing face index, and its boundary faces through
// (bptr, bface). Face quantities needed later (the predicted flux F*, the pressure-correction
// coefficient cf) are written only by the face's owner cell, so nothing is scattered with
// atomics and every result is independent of the launch shape. Sums over a cell's faces keep
// the reference's NumPy order (flowtrace/unstructured/fvmesh.py:254-309): a `scatter` (Cown@ +
// Cnei@ + Cb@) sums owner faces, neighbour faces and boundary faces separately, each in face
// order, then adds them; a net-outflow product (Cdiff@) sums +owner / -neighbour terms
// interleaved in face order; the boundary product Cb@ is added last.
#include "../../include/ftmi.h"
#include "ftmi_common.h"

using namespace ftmi;

namespace {

enum { WALL = 0, INLET = 1, OUTLET = 2, PRESSURE = 3, SYMMETRY = 4 };
enum { T_ADIABATIC = 0, T_FIXED = 1, T_FLUX = 2 };

inline double dmax(double a, double b) { return a > b ? a : b; }
template <typename R>
__device__ inline R fmax0(R a) { return a > R(0) ? a : R(0); }

template <typename R>
__device__ inline R clipd(R x, R lo, R hi) { return x < lo ? lo : (x > hi ? hi : x); }

struct MeshV {  // device view of ftmi_mesh
    int64_t nc, nf, nb;
    const int32_t *P, *N, *Pb, *cptr, *cface, *cnbr, *bptr, *bface;
    const double *w, *alpha, *wl, *S, *d, *Tv, *dPf, *dNf, *dxm;
    const double *Sb, *db, *dbt, *nrm, *alpha_b, *wlb, *Sbm;
    const double *vol, *Ginv;
};

struct BcV {
    const int32_t* kind;
    const int32_t* tkind;
    const double *ubf, *pbf, *tval;  // ubf: (nb, dim)
};

MeshV view(const ftmi_mesh* m) {
    MeshV v;
    v.nc = m->nc;
    v.nf = m->nf;
    v.nb = m->nb;
    v.P = m->P;
    v.N = m->N;
    v.Pb = m->Pb;
    v.cptr = m->cptr;
    v.cface = m->cface;
    v.cnbr = m->cnbr;
    v.bptr = m->bptr;
    v.bface = m->bface;
    v.w = m->w;
    v.alpha = m->alpha;
    v.wl = m->wl;
    v.S = m->S;
    v.d = m->d;
    v.Tv = m->Tvec;
    v.dPf = m->dP_f;
    v.dNf = m->dN_f;
    v.dxm = m->dxm;
    v.Sb = m->Sb;
    v.db = m->db;
    v.dbt = m->db_t;
    v.nrm = m->nb_unit;
    v.alpha_b = m->alpha_b;
    v.wlb = m->wlb;
    v.Sbm = m->Sb_mag;
    v.vol = m->vol;
    v.Ginv = m->Ginv;
    return v;
}

BcV bview(const ftmi_bc* b) { return BcV{b->kind, b->tkind, b->ub_fixed, b->pb_fixed, b->tval}; }

__device__ inline double m_wall(int k) { return k == WALL ? 1.0 : 0.0; }
__device__ inline double m_inlet(int k) { return k == INLET ? 1.0 : 0.0; }
__device__ inline double m_pres(int k) { return (k == OUTLET || k == PRESSURE) ? 1.0 : 0.0; }
__device__ inline double m_sym(int k) { return k == SYMMETRY ? 1.0 : 0.0; }
__device__ inline double m_dirvel(int k) { return (k == WALL || k == INLET || k == SYMMETRY) ? 1.0 : 0.0; }

// ------------------------------------------------------------------ boundary values
// ub, pb of simple.py:305-325 (lagged-gradient reconstruction when gU / gp are given).
template <int DIM>
__global__ void __launch_bounds__(kBlock) k_bvals(MeshV m, BcV bc, const double* __restrict__ U,
                                                  const double* __restrict__ p, const double* __restrict__ gU,
                                                  const double* __restrict__ gp, double* __restrict__ ub,
                                                  double* __restrict__ pb) {
    int64_t lo, hi;
    block_range(m.nb, lo, hi);
    for (int64_t f = lo + threadIdx.x; f < hi; f += kBlock) {
        const int64_t c = m.Pb[f];
        const int kd = bc.kind[f];
        double UP[DIM];
#pragma unroll
        for (int a = 0; a < DIM; ++a) {
            UP[a] = U[a * m.nc + c];
            if (gU) {
                const double* g = gU + (int64_t)a * m.nc * DIM + c * DIM;
                double s = g[0] * m.dbt[f * DIM + 0];
#pragma unroll
                for (int b = 1; b < DIM; ++b) s += g[b] * m.dbt[f * DIM + b];
                UP[a] = UP[a] + s;
            }
        }
        double un = UP[0] * m.nrm[f * DIM];
#pragma unroll
        for (int a = 1; a < DIM; ++a) un += UP[a] * m.nrm[f * DIM + a];
        const double fixed = m_wall(kd) + m_inlet(kd), ms = m_sym(kd), mp = m_pres(kd);
#pragma unroll
        for (int a = 0; a < DIM; ++a) {
            const double sym = UP[a] - un * m.nrm[f * DIM + a];
            ub[a * m.nb + f] = fixed * bc.ubf[f * DIM + a] + ms * sym + mp * UP[a];
        }
        double pP = p[c];
        if (gp) {
            double s = gp[c * DIM] * m.db[f * DIM];
#pragma unroll
            for (int b = 1; b < DIM; ++b) s += gp[c * DIM + b] * m.db[f * DIM + b];
            pP = pP + s;
        }
        pb[f] = mp * bc.pbf[f] + (1.0 - mp) * pP;
    }
}

// Boundary temperatures of simple.py:783-791.
template <int DIM>
__global__ void __launch_bounds__(kBlock) k_tbvals(MeshV m, BcV bc, const double* __restrict__ T,
                                                   const double* __restrict__ gT, const double* __restrict__ Fb,
                                                   double* __restrict__ Tb) {
    int64_t lo, hi;
    block_range(m.nb, lo, hi);
    for (int64_t f = lo + threadIdx.x; f < hi; f += kBlock) {
        const int64_t c = m.Pb[f];
        double TP = T[c];
        if (gT) {
            double s = gT[c * DIM] * m.dbt[f * DIM];
#pragma unroll
            for (int b = 1; b < DIM; ++b) s += gT[c * DIM + b] * m.dbt[f * DIM + b];
            TP = TP + s;
        }
        double t = bc.tkind[f] == T_FIXED ? bc.tval[f] : TP;
        if (bc.kind[f] == OUTLET && Fb[f] < 0.0) t = bc.tval[f];
        Tb[f] = t;
    }
}

// ------------------------------------------------------------------ least-squares gradients
// NF fields at once (fvmesh.py:280-286): r = sum_own + sum_nei + sum_bnd, grad = Ginv r.
struct Fields4 {
    const double* phi[4];
    const double* phib[4];
    double* out[4];
    double bscale[4];  // 0: phib as given; else phib_f = (1 - m_pres_f) * phi_P (pressure correction)
};

template <int DIM, int NF>
__global__ void __launch_bounds__(kBlock) k_grad_ls(MeshV m, Fields4 fl, const int32_t* __restrict__ bkind) {
    int64_t lo, hi;
    block_range(m.nc, lo, hi);
    for (int64_t c = lo + threadIdx.x; c < hi; c += kBlock) {
        double ro[NF][DIM], rn[NF][DIM], rb[NF][DIM], pc[NF];
#pragma unroll
        for (int q = 0; q < NF; ++q) {
            pc[q] = fl.phi[q][c];
#pragma unroll
            for (int a = 0; a < DIM; ++a) ro[q][a] = rn[q][a] = rb[q][a] = 0.0;
        }
        const int32_t e = m.cptr[c + 1];
        for (int32_t k = m.cptr[c]; k < e; ++k) {
            const int32_t code = m.cface[k];
            const int64_t f = code >> 1;
            const bool isN = code & 1;
            const int64_t o = m.cnbr[k];
            const double wl = m.wl[f];
            double dv[DIM];
#pragma unroll
            for (int a = 0; a < DIM; ++a) dv[a] = m.d[f * DIM + a];
#pragma unroll
            for (int q = 0; q < NF; ++q) {
                const double po = fl.phi[q][o];
                const double coef = wl * (isN ? (pc[q] - po) : (po - pc[q]));
                if (isN) {
#pragma unroll
                    for (int a = 0; a < DIM; ++a) rn[q][a] += coef * dv[a];
                } else {
#pragma unroll
                    for (int a = 0; a < DIM; ++a) ro[q][a] += coef * dv[a];
                }
            }
        }
        const int32_t be = m.bptr[c + 1];
        for (int32_t j = m.bptr[c]; j < be; ++j) {
            const int64_t f = m.bface[j];
            const double wlb = m.wlb[f];
#pragma unroll
            for (int q = 0; q < NF; ++q) {
                double pbv;
                if (fl.bscale[q] != 0.0)
                    pbv = (1.0 - m_pres(bkind[f])) * pc[q];
                else
                    pbv = fl.phib[q][f];
                const double coef = wlb * (pbv - pc[q]);
#pragma unroll
                for (int a = 0; a < DIM; ++a) rb[q][a] += coef * m.db[f * DIM + a];
            }
        }
        const double* G = m.Ginv + c * DIM * DIM;
#pragma unroll
        for (int q = 0; q < NF; ++q) {
            double r[DIM];
#pragma unroll
            for (int a = 0; a < DIM; ++a) r[a] = (ro[q][a] + rn[q][a]) + rb[q][a];
#pragma unroll
            for (int a = 0; a < DIM; ++a) {
                double s = G[a * DIM] * r[0];
#pragma unroll
                for (int b = 1; b < DIM; ++b) s += G[a * DIM + b] * r[b];
                fl.out[q][c * DIM + a] = s;
            }
        }
    }
}

// ------------------------------------------------------------------ Green-Gauss pressure force
// (fvmesh.py:303-309): pf = w p_P + (1-w) p_N + (w g_P + (1-w) g_N) . dxm; s = sum +-pf S + sum pb Sb.
template <int DIM>
__global__ void __launch_bounds__(kBlock) k_grad_gg(MeshV m, const double* __restrict__ phi,
                                                    const double* __restrict__ phib, const double* __restrict__ g,
                                                    double* __restrict__ out) {
    int64_t lo, hi;
    block_range(m.nc, lo, hi);
    for (int64_t c = lo + threadIdx.x; c < hi; c += kBlock) {
        double si[DIM], sb[DIM];
#pragma unroll
        for (int a = 0; a < DIM; ++a) si[a] = sb[a] = 0.0;
        const int32_t e = m.cptr[c + 1];
        for (int32_t k = m.cptr[c]; k < e; ++k) {
            const int32_t code = m.cface[k];
            const int64_t f = code >> 1;
            const bool isN = code & 1;
            const int64_t o = m.cnbr[k];
            const int64_t P = isN ? o : c, N = isN ? c : o;
            const double wf = m.w[f], wn = 1.0 - wf;
            double pf = wf * phi[P] + wn * phi[N];
            if (g) {
                double s = (wf * g[P * DIM] + wn * g[N * DIM]) * m.dxm[f * DIM];
#pragma unroll
                for (int a = 1; a < DIM; ++a) s += (wf * g[P * DIM + a] + wn * g[N * DIM + a]) * m.dxm[f * DIM + a];
                pf = pf + s;
            }
#pragma unroll
            for (int a = 0; a < DIM; ++a) {
                const double t = pf * m.S[f * DIM + a];
                si[a] += isN ? -t : t;
            }
        }
        const int32_t be = m.bptr[c + 1];
        for (int32_t j = m.bptr[c]; j < be; ++j) {
            const int64_t f = m.bface[j];
#pragma unroll
            for (int a = 0; a < DIM; ++a) sb[a] += phib[f] * m.Sb[f * DIM + a];
        }
        const double v = m.vol[c];
#pragma unroll
        for (int a = 0; a < DIM; ++a) out[c * DIM + a] = (si[a] + sb[a]) / v;
    }
}

// ------------------------------------------------------------------ transport assembly
// Coefficients (simple.py:340-353), right-hand side (355-374), pre-solve residual sums
// (565-569 / 376-380), relaxed system (570-574 / 382-387) and, for momentum under SIMPLEC,
// the consistent d-coefficient dP = V / (diag_r - sum a_nb) (584-590), for NF transported
// scalars sharing one matrix.
struct AsmArgs {
    const double* F;
    const double* Fb;
    double gamma;
    const double* gam_f;     // per interior face diffusivity (nullptr: gamma everywhere)
    const double* gam_b;     // per boundary face diffusivity (nullptr: gamma)
    const double* extra[3];  // per-cell source added to field q's source; energy: extra[0] (nullptr: none)
    int dir_mode;  // 0: momentum (walls, inlets, symmetry Dirichlet); 1: energy (fixed-T faces)
    const double* phi[3];
    const double* grad[3];
    const double* phib[3];
    const double* gp;  // pressure gradient for the momentum source (nullptr for energy)
    double drive[3];
    int has_drive;
    double relax;
    int second_order;
    int simplec;        // write dP
    double cp;          // energy: boundary heat flux source q = flux * value * |Sb| / cp
    double* dP;
    double* resid;  // resid[2*q] = sum |rhs - A phi|, resid[2*q+1] = sum |diag phi|
    double* x0[3];
    double* cell_resid;  // nullptr, or the per-cell terms of resid: row j of (2 NF, nc)
};

// Arithmetic and outputs in R: double for the FP64 solver, float for the mixed-precision
// transport systems (inputs stay FP64 and are rounded on load; dP stays FP64). In FP32 the
// operator is also written in difference form: ex = diag_r + sum_k off_k, accumulated from
// the face fluxes (each face adds +-F, the diffusion parts cancelling exactly) plus the
// boundary and relaxation terms. The FP32 right-hand side b is then the residual of the
// relaxed system at the current field, r = b_r - A_r phi = rhs - A phi, formed as
// sum_b cb (phi_b - phi_P) - m phi_P - sum off (phi_N - phi_P) + sources (m: the net outflow,
// in FP64), with the differences taken in FP64 and rounded once: FP32 rounding then scales
// with the residual, not with |A| |phi|, whatever the level of phi (a temperature of 300 K
// keeps its millikelvin differences). The FP32 solve finds the correction d of phi.
template <typename R>
struct AsmOut {
    R* diag;
    R* diag_r;
    R* off;
    R* b[3];
    R* inv_r;
    R* ex;  // FP32 only (difference form)
};

template <int DIM, int NF, typename R>
__global__ void __launch_bounds__(kBlock) k_assemble(MeshV m, BcV bc, AsmArgs A, AsmOut<R> O, RedBuf rb) {
    int64_t lo, hi;
    block_range(m.nc, lo, hi);
    R acc[2 * NF];
#pragma unroll
    for (int j = 0; j < 2 * NF; ++j) acc[j] = R(0);
    const R gamma = (R)A.gamma;
    for (int64_t c = lo + threadIdx.x; c < hi; c += kBlock) {
        constexpr bool kDiff = sizeof(R) == 4;
        R dg_o = R(0), dg_n = R(0), dg_b = R(0);   // diagonal: owner, neighbour, boundary sums
        R nb_o = R(0), nb_n = R(0);                // sum of -a_nb (SIMPLEC)
        R exs = R(0);                              // kDiff: sum of +-F over the faces
        // rt, rd: Cdiff products (owner +, neighbour -, interleaved in face order)
        R rb_[NF], rt[NF], rd[NF], ax[NF], pc[NF];
        double pc64[NF];
        double m64 = 0.0;  // kDiff: the cell's net outflow, sum of +-F and Fb (FP64)
#pragma unroll
        for (int q = 0; q < NF; ++q) {
            rb_[q] = rt[q] = rd[q] = ax[q] = R(0);
            pc64[q] = A.phi[q][c];
            pc[q] = (R)pc64[q];
        }
        const int32_t e = m.cptr[c + 1];
        for (int32_t k = m.cptr[c]; k < e; ++k) {
            const int32_t code = m.cface[k];
            const int64_t f = code >> 1;
            const bool isN = code & 1;
            const int64_t o = m.cnbr[k];
            const R F = (R)A.F[f];
            const R gf = A.gam_f ? (R)A.gam_f[f] : gamma;
            const R D = gf * (R)m.alpha[f];
            const R aPN = -(D + fmax0(-F));
            const R aNP = -(D + fmax0(F));
            const R a_here = isN ? aNP : aPN;
            O.off[k] = a_here;
            if (kDiff) {
                exs += isN ? -F : F;
                m64 += isN ? -A.F[f] : A.F[f];
            }
            if (isN) {
                dg_n += D + fmax0(-F);
                nb_n += -aNP;
            } else {
                dg_o += D + fmax0(F);
                nb_o += -aPN;
            }
            const int64_t P = isN ? o : c, N = isN ? c : o;
            const R wf = (R)m.w[f], wn = R(1) - wf;
#pragma unroll
            for (int q = 0; q < NF; ++q) {
                const double* g = A.grad[q];
                const double po64 = A.phi[q][o];
                const R po = (R)po64;
                // kDiff: the neighbour difference is taken in FP64 and rounded once
                ax[q] += kDiff ? a_here * (R)(po64 - pc64[q]) : a_here * po;
                // non-orthogonal diffusion gamma (grad phi)_f . T
                R gt = (wf * (R)g[P * DIM] + wn * (R)g[N * DIM]) * (R)m.Tv[f * DIM];
#pragma unroll
                for (int a = 1; a < DIM; ++a)
                    gt += (wf * (R)g[P * DIM + a] + wn * (R)g[N * DIM + a]) * (R)m.Tv[f * DIM + a];
                const R tn = gf * gt;
                R dc = R(0);
                if (A.second_order) {
                    const R phP = isN ? po : pc[q], phN = isN ? pc[q] : po;
                    R sP = (R)g[P * DIM] * (R)m.dPf[f * DIM], sN = (R)g[N * DIM] * (R)m.dNf[f * DIM];
#pragma unroll
                    for (int a = 1; a < DIM; ++a) {
                        sP += (R)g[P * DIM + a] * (R)m.dPf[f * DIM + a];
                        sN += (R)g[N * DIM + a] * (R)m.dNf[f * DIM + a];
                    }
                    const bool up = F >= R(0);
                    const R hoP = phP + sP, hoN = phN + sN;
                    const R lo_ = phP < phN ? phP : phN, hi_ = phP > phN ? phP : phN;
                    const R ho = clipd(up ? hoP : hoN, lo_, hi_);
                    dc = F * (ho - (up ? phP : phN));
                }
                rt[q] += isN ? -tn : tn;
                rd[q] += isN ? -dc : dc;
            }
        }
        const int32_t be = m.bptr[c + 1];
        R qin = R(0);
        for (int32_t j = m.bptr[c]; j < be; ++j) {
            const int64_t f = m.bface[j];
            const int kd = bc.kind[f];
            const R Fb = (R)A.Fb[f];
            R dir;
            if (A.dir_mode == 0)
                dir = (R)m_dirvel(kd);
            else
                dir = bc.tkind[f] == T_FIXED ? R(1) : R(0);
            const R Db = (A.gam_b ? (R)A.gam_b[f] : gamma) * (R)m.alpha_b[f] * dir;
            dg_b += Db + fmax0(Fb);
            const R cb = Db + fmax0(-Fb);
            if (kDiff) {
                // difference form of the boundary source: cb (phi_b - phi_P), the rest of
                // cb phi_b - (Db + max(Fb, 0)) phi_P being -Fb phi_P, which joins the net
                // outflow term
                m64 += A.Fb[f];
#pragma unroll
                for (int q = 0; q < NF; ++q) rb_[q] += cb * (R)(A.phib[q][f] - pc64[q]);
            } else {
#pragma unroll
                for (int q = 0; q < NF; ++q) rb_[q] += cb * (R)A.phib[q][f];
            }
            if (A.dir_mode == 1)
                qin += (bc.tkind[f] == T_FLUX ? R(1) : R(0)) * (R)bc.tval[f] * (R)m.Sbm[f] / (R)A.cp;
        }
        const R relax = (R)A.relax;
        const R diag = (dg_o + dg_n) + dg_b;
        const R diag_r = diag / relax;
        O.diag[c] = diag;
        O.diag_r[c] = diag_r;
        if (O.inv_r) O.inv_r[c] = R(1) / diag_r;
        const R vol = (R)m.vol[c];
        const R crel = (R(1) - relax) / relax;
        const R ex_u = exs + dg_b;  // kDiff: diag + sum_k off_k of the unrelaxed operator
        if (kDiff) O.ex[c] = ex_u + crel * diag;
#pragma unroll
        for (int q = 0; q < NF; ++q) {
            // rhs = Cb@(bnd) + Cdiff@(nonorth) - Cdiff@(deferred) + extra
            R rhs = (rb_[q] + rt[q]) - rd[q];
            if (A.dir_mode == 0) {
                R src = -(R)A.gp[c * DIM + q] * vol;
                if (A.has_drive) src = src + vol * (R)A.drive[q];
                if (A.extra[q]) src = src + (R)A.extra[q][c];
                rhs = rhs + src;
            } else {
                rhs = rhs + (A.extra[0] ? qin + (R)A.extra[0][c] : qin);  // energy source after the flux walls
            }
            // kDiff: rhs holds the boundary source in difference form, so the residual is
            // rhs - m phi_P - sum off (phi_N - phi_P), every term small where the field is
            // smooth whatever its level (no cancellation of a_P phi against b in FP32); the
            // FP32 system's right-hand side is that residual, for the correction of phi
            const R resid = kDiff ? (rhs - (R)m64 * pc[q]) - ax[q] : rhs - (diag * pc[q] + ax[q]);
            const R r_abs = fabs(resid), d_abs = fabs(diag * pc[q]);
            acc[2 * q] += r_abs;
            acc[2 * q + 1] += d_abs;
            if (A.cell_resid) {
                A.cell_resid[(2 * q) * m.nc + c] = (double)r_abs;
                A.cell_resid[(2 * q + 1) * m.nc + c] = (double)d_abs;
            }
            O.b[q][c] = kDiff ? resid : rhs + crel * diag * pc[q];
            if (A.x0[q]) A.x0[q][c] = A.phi[q][c];
        }
        if (A.simplec >= 0 && A.dP) {
            R denom = diag_r;
            if (A.simplec == 1) {
                denom = diag_r - (nb_o + nb_n);
                denom = denom > R(1e-12) * diag_r ? denom : diag_r;
            }
            A.dP[c] = (double)(vol / denom);
        }
    }
    block_partials<2 * NF>(acc, rb.partials);
}

// ------------------------------------------------------------------ Rhie-Chow fluxes + pressure system
// simple.py:590-621: F*, boundary F*, cf, cb, A_p (diagonal + -cf), mass imbalance, sum |mass|.
template <int DIM>
__global__ void __launch_bounds__(kBlock) k_rhie_chow(MeshV m, BcV bc, double rho, const double* __restrict__ Us,
                                                      const double* __restrict__ p, const double* __restrict__ gp,
                                                      const double* __restrict__ dP, double* __restrict__ Fst,
                                                      double* __restrict__ cf, double* __restrict__ Fbst,
                                                      double* __restrict__ cbv, double* __restrict__ diag_p,
                                                      double* __restrict__ offp, double* __restrict__ mass,
                                                      double* __restrict__ neg_mass, double* cont_out, RedBuf rb) {
    int64_t lo, hi;
    block_range(m.nc, lo, hi);
    double acc[1] = {0.0};
    for (int64_t c = lo + threadIdx.x; c < hi; c += kBlock) {
        double mi_ = 0.0, mb = 0.0, dgo = 0.0, dgn = 0.0, dgb = 0.0;
        const int32_t e = m.cptr[c + 1];
        for (int32_t k = m.cptr[c]; k < e; ++k) {
            const int32_t code = m.cface[k];
            const int64_t f = code >> 1;
            const bool isN = code & 1;
            const int64_t o = m.cnbr[k];
            const int64_t P = isN ? o : c, N = isN ? c : o;
            const double wf = m.w[f], wn = 1.0 - wf;
            const double dF = wf * dP[P] + wn * dP[N];
            double us = (wf * Us[P] + wn * Us[N]) * m.S[f * DIM];
            double gd = (wf * gp[P * DIM] + wn * gp[N * DIM]) * m.d[f * DIM];
#pragma unroll
            for (int a = 1; a < DIM; ++a) {
                us += (wf * Us[a * m.nc + P] + wn * Us[a * m.nc + N]) * m.S[f * DIM + a];
                gd += (wf * gp[P * DIM + a] + wn * gp[N * DIM + a]) * m.d[f * DIM + a];
            }
            const double al = m.alpha[f];
            const double Fs = rho * (us - dF * al * (p[N] - p[P] - gd));
            const double cff = rho * dF * al;
            offp[k] = -cff;
            mi_ += isN ? -Fs : Fs;
            if (isN) {
                dgn += cff;
            } else {
                dgo += cff;
                Fst[f] = Fs;
                cf[f] = cff;
            }
        }
        const int32_t be = m.bptr[c + 1];
        for (int32_t j = m.bptr[c]; j < be; ++j) {
            const int64_t f = m.bface[j];
            const int kd = bc.kind[f];
            const double mi = m_inlet(kd), mp = m_pres(kd);
            double fin = bc.ubf[f * DIM] * m.Sb[f * DIM];
            double us = Us[c] * m.Sb[f * DIM];
            double gd = gp[c * DIM] * m.db[f * DIM];
#pragma unroll
            for (int a = 1; a < DIM; ++a) {
                fin += bc.ubf[f * DIM + a] * m.Sb[f * DIM + a];
                us += Us[a * m.nc + c] * m.Sb[f * DIM + a];
                gd += gp[c * DIM + a] * m.db[f * DIM + a];
            }
            const double Fb_in = rho * fin * mi;
            const double Fb_pr = rho * (us - dP[c] * m.alpha_b[f] * (bc.pbf[f] - p[c] - gd));
            const double Fbs = Fb_in + Fb_pr * mp;
            const double cbb = rho * dP[c] * m.alpha_b[f] * mp;
            Fbst[f] = Fbs;
            cbv[f] = cbb;
            mb += Fbs;
            dgb += cbb;
        }
        const double ms = mi_ + mb;
        mass[c] = ms;
        if (neg_mass) neg_mass[c] = -ms;
        diag_p[c] = (dgo + dgn) + dgb;
        acc[0] += fabs(ms);
    }
    block_partials<1>(acc, rb.partials);
}

// ------------------------------------------------------------------ corrections
// simple.py:639-646: F = F* + cf (p'_P - p'_N), Fb = Fb* + cb p'_P, p += relax_p p',
// grad p' (LS, p'_b = (1 - m_pres) p'_P), U = U* - dP grad p'; sums of p, |F|, |Fb|.
template <int DIM>
__global__ void __launch_bounds__(kBlock) k_correct(MeshV m, BcV bc, const double* __restrict__ pp,
                                                    const double* __restrict__ Us, const double* __restrict__ dP,
                                                    const double* __restrict__ Fst, const double* __restrict__ cf,
                                                    const double* __restrict__ Fbst, const double* __restrict__ cbv,
                                                    double relax_p, double* __restrict__ U, double* __restrict__ p,
                                                    double* __restrict__ F, double* __restrict__ Fb,
                                                    double* __restrict__ gpp_out, double* sums_out, RedBuf rb) {
    int64_t lo, hi;
    block_range(m.nc, lo, hi);
    double acc[3] = {0.0, 0.0, 0.0};
    for (int64_t c = lo + threadIdx.x; c < hi; c += kBlock) {
        const double pc = pp[c];
        double ro[DIM], rn[DIM], rbv[DIM];
#pragma unroll
        for (int a = 0; a < DIM; ++a) ro[a] = rn[a] = rbv[a] = 0.0;
        const int32_t e = m.cptr[c + 1];
        for (int32_t k = m.cptr[c]; k < e; ++k) {
            const int32_t code = m.cface[k];
            const int64_t f = code >> 1;
            const bool isN = code & 1;
            const double po = pp[m.cnbr[k]];
            const double coef = m.wl[f] * (isN ? (pc - po) : (po - pc));
            if (isN) {
#pragma unroll
                for (int a = 0; a < DIM; ++a) rn[a] += coef * m.d[f * DIM + a];
            } else {
#pragma unroll
                for (int a = 0; a < DIM; ++a) ro[a] += coef * m.d[f * DIM + a];
                const double Ff = Fst[f] + cf[f] * (pc - po);
                F[f] = Ff;
                acc[1] += fabs(Ff);
            }
        }
        const int32_t be = m.bptr[c + 1];
        for (int32_t j = m.bptr[c]; j < be; ++j) {
            const int64_t f = m.bface[j];
            const double ppb = (1.0 - m_pres(bc.kind[f])) * pc;
            const double coef = m.wlb[f] * (ppb - pc);
#pragma unroll
            for (int a = 0; a < DIM; ++a) rbv[a] += coef * m.db[f * DIM + a];
            const double Fbf = Fbst[f] + cbv[f] * pc;
            Fb[f] = Fbf;
            acc[2] += fabs(Fbf);
        }
        const double* G = m.Ginv + c * DIM * DIM;
        double r[DIM];
#pragma unroll
        for (int a = 0; a < DIM; ++a) r[a] = (ro[a] + rn[a]) + rbv[a];
        const double dPc = dP[c];
#pragma unroll
        for (int a = 0; a < DIM; ++a) {
            double s = G[a * DIM] * r[0];
#pragma unroll
            for (int b = 1; b < DIM; ++b) s += G[a * DIM + b] * r[b];
            if (gpp_out) gpp_out[c * DIM + a] = s;
            U[a * m.nc + c] = Us[a * m.nc + c] - dPc * s;
        }
        const double pn = p[c] + relax_p * pc;
        p[c] = pn;
        acc[0] += pn;
    }
    block_partials<3>(acc, rb.partials);
}

__global__ void __launch_bounds__(kBlock) k_sub_mean(int64_t n, double* __restrict__ p, const double* sum) {
    const double mean = *sum / (double)n;
    int64_t lo, hi;
    block_range(n, lo, hi);
    for (int64_t i = lo + threadIdx.x; i < hi; i += kBlock) p[i] = p[i] - mean;
}

// The initial face fluxes of simple.py:110-112: F = rho (U_f . S), Fb = rho (ub . Sb) (m_inlet + m_pres).
template <int DIM>
__global__ void __launch_bounds__(kBlock) k_init_flux(MeshV m, BcV bc, double rho, const double* __restrict__ U,
                                                      const double* __restrict__ ub, double* __restrict__ F,
                                                      double* __restrict__ Fb) {
    int64_t lo, hi;
    block_range(m.nf + m.nb, lo, hi);
    for (int64_t t = lo + threadIdx.x; t < hi; t += kBlock) {
        if (t < m.nf) {
            const int64_t f = t, P = m.P[f], N = m.N[f];
            const double wf = m.w[f], wn = 1.0 - wf;
            double s = (wf * U[P] + wn * U[N]) * m.S[f * DIM];
#pragma unroll
            for (int a = 1; a < DIM; ++a) s += (wf * U[a * m.nc + P] + wn * U[a * m.nc + N]) * m.S[f * DIM + a];
            F[f] = rho * s;
        } else {
            const int64_t f = t - m.nf;
            double s = ub[f] * m.Sb[f * DIM];
#pragma unroll
            for (int a = 1; a < DIM; ++a) s += ub[a * m.nb + f] * m.Sb[f * DIM + a];
            const int kd = bc.kind[f];
            Fb[f] = rho * s * (m_inlet(kd) + m_pres(kd));
        }
    }
}

// sum |x| over n values (used for the flux scale of freshly initialised fields)
__global__ void __launch_bounds__(kBlock) k_abs_sum(int64_t n, const double* __restrict__ x, double* out, RedBuf rb) {
    int64_t lo, hi;
    block_range(n, lo, hi);
    double acc[1] = {0.0};
    for (int64_t i = lo + threadIdx.x; i < hi; i += kBlock) acc[0] += fabs(x[i]);
    block_partials<1>(acc, rb.partials);
}

RedBuf red(const ftmi_red* r) { return RedBuf{r->partials, r->counter}; }

}  // namespace

#define DISPATCH_DIM(dim, ...)                              \
    do {                                                    \
        if ((dim) == 2) {                                   \
            constexpr int D = 2;                            \
            __VA_ARGS__;                                    \
        } else if ((dim) == 3) {                            \
            constexpr int D = 3;                            \
            __VA_ARGS__;                                    \
        } else {                                            \
            set_error_msg("dim must be 2 or 3");            \
            return -1;                                      \
        }                                                   \
    } while (0)

extern "C" {

int ftmi_simplec_bvals(const ftmi_mesh* mesh, const ftmi_bc* bc, const double* U, const double* p, const double* gU,
                       const double* gp, double* ub, double* pb, void* stream) {
    MeshV m = view(mesh);
    if (m.nb == 0) return 0;
    DISPATCH_DIM(mesh->dim, (k_bvals<D><<<grid_for(m.nb), kBlock, 0, (hipStream_t)stream>>>(m, bview(bc), U, p, gU,
                                                                                             gp, ub, pb)));
    FTMI_LAUNCH_CHECK("ftmi_simplec_bvals");
    return 0;
}

int ftmi_energy_bvals(const ftmi_mesh* mesh, const ftmi_bc* bc, const double* T, const double* gT, const double* Fb,
                      double* Tb, void* stream) {
    MeshV m = view(mesh);
    if (m.nb == 0) return 0;
    DISPATCH_DIM(mesh->dim, (k_tbvals<D><<<grid_for(m.nb), kBlock, 0, (hipStream_t)stream>>>(m, bview(bc), T, gT, Fb,
                                                                                              Tb)));
    FTMI_LAUNCH_CHECK("ftmi_energy_bvals");
    return 0;
}

int ftmi_ls_gradients(const ftmi_mesh* mesh, int nfields, const double* const* phi, const double* const* phib,
                      double* const* out, const int32_t* bkind_for_pprime, void* stream) {
    MeshV m = view(mesh);
    Fields4 fl{};
    if (nfields < 1 || nfields > 4) {
        set_error_msg("ftmi_ls_gradients: 1..4 fields");
        return -1;
    }
    for (int q = 0; q < nfields; ++q) {
        fl.phi[q] = phi[q];
        fl.phib[q] = phib ? phib[q] : nullptr;
        fl.out[q] = out[q];
        fl.bscale[q] = (phib == nullptr || phib[q] == nullptr) ? 1.0 : 0.0;
    }
    hipStream_t s = (hipStream_t)stream;
    const int g = grid_for(m.nc);
#define LSCASE(NF) \
    case NF:       \
        DISPATCH_DIM(mesh->dim, (k_grad_ls<D, NF><<<g, kBlock, 0, s>>>(m, fl, bkind_for_pprime))); \
        break;
    switch (nfields) {
        LSCASE(1)
        LSCASE(2)
        LSCASE(3)
        LSCASE(4)
    }
#undef LSCASE
    FTMI_LAUNCH_CHECK("ftmi_ls_gradients");
    return 0;
}

int ftmi_gg_gradient_fused(const ftmi_mesh* mesh, const double* phi, const double* phib, const double* grad,
                           double* out, void* stream) {
    MeshV m = view(mesh);
    DISPATCH_DIM(mesh->dim, (k_grad_gg<D><<<grid_for(m.nc), kBlock, 0, (hipStream_t)stream>>>(m, phi, phib, grad,
                                                                                                out)));
    FTMI_LAUNCH_CHECK("ftmi_gg_gradient_fused");
    return 0;
}

}  // extern "C"

namespace {
template <typename R>
void launch_assemble(int dim, bool energy, int g, hipStream_t s, const MeshV& m, const BcV& b, const AsmArgs& A,
                     const AsmOut<R>& O, RedBuf rb) {
    if (energy) {
        if (dim == 2)
            k_assemble<2, 1, R><<<g, kBlock, 0, s>>>(m, b, A, O, rb);
        else
            k_assemble<3, 1, R><<<g, kBlock, 0, s>>>(m, b, A, O, rb);
        fin_store<2><<<1, kBlock, 0, s>>>(rb.partials, g, A.resid, 1.0);
    } else if (dim == 2) {
        k_assemble<2, 2, R><<<g, kBlock, 0, s>>>(m, b, A, O, rb);
        fin_store<4><<<1, kBlock, 0, s>>>(rb.partials, g, A.resid, 1.0);
    } else {
        k_assemble<3, 3, R><<<g, kBlock, 0, s>>>(m, b, A, O, rb);
        fin_store<6><<<1, kBlock, 0, s>>>(rb.partials, g, A.resid, 1.0);
    }
}
}  // namespace

extern "C" {

int ftmi_transport_assemble(const ftmi_mesh* mesh, const ftmi_bc* bc, const ftmi_assemble_args* a,
                            const ftmi_red* rbuf, void* stream) {
    MeshV m = view(mesh);
    AsmArgs A{};
    A.F = a->F;
    A.Fb = a->Fb;
    A.gamma = a->gamma;
    A.gam_f = a->gam_f;
    A.gam_b = a->gam_b;
    for (int q = 0; q < 3; ++q) A.extra[q] = a->extra[q];
    A.dir_mode = a->energy ? 1 : 0;
    const int nfld = a->nfields;
    for (int q = 0; q < nfld && q < 3; ++q) {
        A.phi[q] = a->phi[q];
        A.grad[q] = a->grad[q];
        A.phib[q] = a->phib[q];
    }
    A.gp = a->gp;
    for (int q = 0; q < 3; ++q) A.drive[q] = a->drive[q];
    A.has_drive = a->has_drive;
    A.relax = a->relax;
    A.second_order = a->second_order;
    A.simplec = a->simplec;
    A.cp = a->cp;
    A.dP = a->dP;
    A.resid = a->resid;
    A.cell_resid = a->cell_resid;
    for (int q = 0; q < 3; ++q) A.x0[q] = q < nfld ? a->x0[q] : nullptr;
    hipStream_t s = (hipStream_t)stream;
    const int g = grid_part(m.nc);
    RedBuf rb = red(rbuf);
    if (a->energy ? nfld != 1 : nfld != mesh->dim) {
        set_error_msg(a->energy ? "energy assembly takes one field" : "momentum assembly takes dim fields");
        return -1;
    }
    if (a->out_fp32) {
        if (!a->ex32) {
            set_error_msg("ftmi_transport_assemble: out_fp32 needs ex32");
            return -1;
        }
        AsmOut<float> O{a->diag32, a->diag_r32, a->off32, {a->b32[0], a->b32[1], a->b32[2]}, a->inv_r32, a->ex32};
        launch_assemble<float>(mesh->dim, a->energy != 0, g, s, m, bview(bc), A, O, rb);
    } else {
        AsmOut<double> O{a->diag, a->diag_r, a->off, {a->b[0], a->b[1], a->b[2]}, a->inv_r, nullptr};
        launch_assemble<double>(mesh->dim, a->energy != 0, g, s, m, bview(bc), A, O, rb);
    }
    FTMI_LAUNCH_CHECK("ftmi_transport_assemble");
    return 0;
}

int ftmi_rhie_chow(const ftmi_mesh* mesh, const ftmi_bc* bc, double rho, const double* Ustar, const double* p,
                   const double* gp, const double* dP, double* Fst, double* cf, double* Fbst, double* cb,
                   double* diag_p, double* offp, double* mass, double* neg_mass, double* cont_out,
                   const ftmi_red* rbuf, void* stream) {
    MeshV m = view(mesh);
    const int g = grid_part(m.nc);
    DISPATCH_DIM(mesh->dim, (k_rhie_chow<D><<<g, kBlock, 0, (hipStream_t)stream>>>(
                                m, bview(bc), rho, Ustar, p, gp, dP, Fst, cf, Fbst, cb, diag_p, offp, mass, neg_mass,
                                cont_out, red(rbuf))));
    fin_store<1><<<1, kBlock, 0, (hipStream_t)stream>>>(rbuf->partials, g, cont_out, 1.0);
    FTMI_LAUNCH_CHECK("ftmi_rhie_chow");
    return 0;
}

int ftmi_correct(const ftmi_mesh* mesh, const ftmi_bc* bc, const double* pp, const double* Ustar, const double* dP,
                 const double* Fst, const double* cf, const double* Fbst, const double* cb, double relax_p,
                 int remove_mean, double* U, double* p, double* F, double* Fb, double* gpp_out, double* sums_out,
                 const ftmi_red* rbuf, void* stream) {
    MeshV m = view(mesh);
    hipStream_t s = (hipStream_t)stream;
    const int g = grid_part(m.nc);
    DISPATCH_DIM(mesh->dim, (k_correct<D><<<g, kBlock, 0, s>>>(m, bview(bc), pp, Ustar, dP, Fst, cf, Fbst, cb, relax_p,
                                                                U, p, F, Fb, gpp_out, sums_out, red(rbuf))));
    fin_store<3><<<1, kBlock, 0, s>>>(rbuf->partials, g, sums_out, 1.0);
    if (remove_mean) k_sub_mean<<<grid_for(m.nc), kBlock, 0, s>>>(m.nc, p, &sums_out[0]);
    FTMI_LAUNCH_CHECK("ftmi_correct");
    return 0;
}

int ftmi_init_flux(const ftmi_mesh* mesh, const ftmi_bc* bc, double rho, const double* U, const double* ub,
                   double* F, double* Fb, void* stream) {
    MeshV m = view(mesh);
    DISPATCH_DIM(mesh->dim, (k_init_flux<D><<<grid_for(m.nf + m.nb), kBlock, 0, (hipStream_t)stream>>>(
                                m, bview(bc), rho, U, ub, F, Fb)));
    FTMI_LAUNCH_CHECK("ftmi_init_flux");
    return 0;
}

int ftmi_abs_sum(int64_t n, const double* x, double* out, const ftmi_red* rbuf, void* stream) {
    k_abs_sum<<<grid_part(n), kBlock, 0, (hipStream_t)stream>>>(n, x, out, red(rbuf));
    fin_store<1><<<1, kBlock, 0, (hipStream_t)stream>>>(rbuf->partials, grid_part(n), out, 1.0);
    FTMI_LAUNCH_CHECK("ftmi_abs_sum");
    return 0;
}

}  // extern "C"
