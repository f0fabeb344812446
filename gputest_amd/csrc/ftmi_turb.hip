// k-omega SST (Menter 2003) of the outer iteration (flowtrace/unstructured/turbulence.py:116-437,
// wired in by flowtrace/unstructured/simple.py:546-548, 662-669, 781-782).
//
// One SST update after the energy step: boundary values of k and omega (fixed at inlets and at
// open faces with backflow), their LS gradients (ftmi_ls_gradients), one cell pass for the strain
// rate, the blending function F1 and the sources and diffusivities of both transports (all from
// the fields before the update), then per transport an LDU assembly with a cell-varying
// diffusivity, the near-wall omega penalty, BiCGSTAB, and finally the floor and mu_t from the new
// fields. Every expression keeps the reference's NumPy evaluation order, so the arithmetic is the
// reference's up to the last bit of exp/log/tanh/pow; sums over a cell's faces run in face order
// like the reference's incidence products, reductions are fixed-order.
#include "../../include/ftmi.h"
#include "ftmi_common.h"

using namespace ftmi;

namespace {

enum { WALL = 0, INLET = 1, OUTLET = 2, PRESSURE = 3, SYMMETRY = 4 };

// model constants (turbulence.py:50-59, 352-360)
constexpr double kKappa = 0.41, kEWall = 9.793;
constexpr double kSk1 = 0.85, kSw1 = 0.5, kB1 = 0.075, kSk2 = 1.0, kSw2 = 0.856, kB2 = 0.0828;
constexpr double kBstar = 0.09, kA1 = 0.31, kPrT = 0.85;

__device__ inline double dmax(double a, double b) { return a > b ? a : b; }
__device__ inline double dmin(double a, double b) { return a < b ? a : b; }
__device__ inline double mix(double F1, double a, double b) { return F1 * a + (1.0 - F1) * b; }

// Boundary values of k and omega (turbulence.py:263-269): the inflow value on inlets and on
// outlet/pressure faces with backflow (Fb < 0), the owner cell's value elsewhere.
__global__ void __launch_bounds__(kBlock) k_sst_bvals(int64_t nb, const int32_t* __restrict__ Pb,
                                                      const int32_t* __restrict__ kind, const double* __restrict__ Fb,
                                                      const double* __restrict__ k, const double* __restrict__ w,
                                                      const double* __restrict__ kin, const double* __restrict__ win,
                                                      double* __restrict__ kb, double* __restrict__ wb,
                                                      double* __restrict__ fixed) {
    int64_t lo, hi;
    block_range(nb, lo, hi);
    for (int64_t f = lo + threadIdx.x; f < hi; f += kBlock) {
        const int kd = kind[f];
        const bool fx = kd == INLET || ((kd == OUTLET || kd == PRESSURE) && Fb[f] < 0.0);
        const int64_t c = Pb[f];
        kb[f] = fx ? kin[f] : k[c];
        wb[f] = fx ? win[f] : w[c];
        fixed[f] = fx ? 1.0 : 0.0;
    }
}

// Strain rate, F1, grad k . grad omega, and the sources (times the cell volume) and cell
// diffusivities of the k and omega transports (turbulence.py:99-113, 367-380, 419-433).
// gU: (DIM, nc, DIM), gU[i][c][j] = d u_i / d x_j; gk, gw: (nc, DIM).
template <int DIM>
__global__ void __launch_bounds__(kBlock) k_sst_cell(int64_t n, const double* __restrict__ gU,
                                                     const double* __restrict__ gk, const double* __restrict__ gw,
                                                     const double* __restrict__ k, const double* __restrict__ w,
                                                     const double* __restrict__ mut, const double* __restrict__ y,
                                                     const double* __restrict__ vol, ftmi_sst_coef C,
                                                     double* __restrict__ S, double* __restrict__ gam_k,
                                                     double* __restrict__ sck, double* __restrict__ spk,
                                                     double* __restrict__ gam_w, double* __restrict__ scw,
                                                     double* __restrict__ spw) {
    int64_t lo, hi;
    block_range(n, lo, hi);
    for (int64_t c = lo + threadIdx.x; c < hi; c += kBlock) {
        double s2 = 0.0;
#pragma unroll
        for (int i = 0; i < DIM; ++i)
#pragma unroll
            for (int j = 0; j < DIM; ++j) {
                const double sij = 0.5 * (gU[(i * n + c) * DIM + j] + gU[(j * n + c) * DIM + i]);
                s2 = s2 + sij * sij;
            }
        const double smag = sqrt(2.0 * s2);  // |S| = sqrt(2 S_ij S_ij)
        double kw = gk[c * DIM] * gw[c * DIM];
#pragma unroll
        for (int a = 1; a < DIM; ++a) kw += gk[c * DIM + a] * gw[c * DIM + a];
        const double kc = k[c], wc = w[c], yc = y[c], mt = mut[c], v = vol[c];
        // blending (turbulence.py:367-380)
        const double cd = dmax(C.two_rho_sw2 * kw / wc, 1e-10);
        const double sk = sqrt(kc);
        const double visc = C.nu500 / (yc * yc * wc);
        const double a = dmax(sk / (kBstar * wc * yc), visc);
        const double arg1 = dmin(a, C.four_rho_sw2 * kc / (cd * yc * yc));
        const double F1 = tanh(pow(arg1, 4.0));
        // sources (turbulence.py:419-433)
        const double S2 = smag * smag;
        const double Pk = dmin(mt * S2, C.ten_bstar_rho * kc * wc);
        gam_k[c] = C.mu + mix(F1, kSk1, kSk2) * mt;
        sck[c] = Pk * v;
        spk[c] = C.bstar_rho * wc * v;
        const double cross = (1.0 - F1) * 2.0 * C.rho * kSw2 * kw / wc;
        const double Scw = mix(F1, C.g1, C.g2) * C.rho * S2 + dmax(cross, 0.0);
        const double Spw = mix(F1, kB1, kB2) * C.rho * wc + dmax(-cross, 0.0) / wc;
        gam_w[c] = C.mu + mix(F1, kSw1, kSw2) * mt;
        scw[c] = Scw * v;
        spw[c] = Spw * v;
        S[c] = smag;
    }
}

// Transport of one SST scalar (turbulence.py:271-304 with simple.py:340-374): upwind convection,
// diffusion with the face interpolation of the cell diffusivity gam, Dirichlet where `fixed`
// (boundary), sources Sc (rhs) and Sp (diagonal), relaxation. The residual sums skip cells at the
// floor and fixed cells. Writes diag (unrelaxed, before any fixed-cell penalty), rhs, the relaxed
// diag_r / b / inv_r, x0 = phi, and the workgroup partials: (sum |rhs - A phi|, sum |diag phi|)
// at sums[2 b + j] and max diag at dmaxp[b].
template <int DIM>
__global__ void __launch_bounds__(kBlock) k_sst_assemble(ftmi_mesh m, const double* __restrict__ F,
                                                         const double* __restrict__ Fb,
                                                         const double* __restrict__ gam,
                                                         const double* __restrict__ phi,
                                                         const double* __restrict__ grad,
                                                         const double* __restrict__ phib,
                                                         const double* __restrict__ fixed,
                                                         const double* __restrict__ Sc,
                                                         const double* __restrict__ Sp,
                                                         const uint8_t* __restrict__ held, double floor_v,
                                                         double relax, double* __restrict__ diag_o,
                                                         double* __restrict__ rhs_o, double* __restrict__ diag_r,
                                                         double* __restrict__ off, double* __restrict__ b,
                                                         double* __restrict__ inv_r, double* __restrict__ x0,
                                                         double* __restrict__ sums, double* __restrict__ dmaxp) {
    __shared__ double red[kBlock];
    int64_t lo, hi;
    block_range(m.nc, lo, hi);
    double acc[2] = {0.0, 0.0};
    double mx = -INFINITY;
    const double crel = (1.0 - relax) / relax;
    for (int64_t c = lo + threadIdx.x; c < hi; c += kBlock) {
        double dg_o = 0.0, dg_n = 0.0, dg_b = 0.0, rt = 0.0, rbn = 0.0, ax = 0.0;
        const double pc = phi[c];
        const int32_t e = m.cptr[c + 1];
        for (int32_t h = m.cptr[c]; h < e; ++h) {
            const int32_t code = m.cface[h];
            const int64_t f = code >> 1;
            const bool isN = code & 1;
            const int64_t o = m.cnbr[h];
            const int64_t P = isN ? o : c, N = isN ? c : o;
            const double wf = m.w[f], wn = 1.0 - wf;
            const double gf = wf * gam[P] + wn * gam[N];
            const double Ff = F[f];
            const double D = gf * m.alpha[f];
            const double aPN = -(D + dmax(-Ff, 0.0)), aNP = -(D + dmax(Ff, 0.0));
            const double a_here = isN ? aNP : aPN;
            off[h] = a_here;
            if (isN)
                dg_n += D + dmax(-Ff, 0.0);
            else
                dg_o += D + dmax(Ff, 0.0);
            ax += a_here * phi[o];
            double gt = (wf * grad[P * DIM] + wn * grad[N * DIM]) * m.Tvec[f * DIM];
#pragma unroll
            for (int a = 1; a < DIM; ++a) gt += (wf * grad[P * DIM + a] + wn * grad[N * DIM + a]) * m.Tvec[f * DIM + a];
            const double tn = gf * gt;
            rt += isN ? -tn : tn;
        }
        const int32_t be = m.bptr[c + 1];
        for (int32_t j = m.bptr[c]; j < be; ++j) {
            const int64_t f = m.bface[j];
            const double Fbf = Fb[f];
            const double Db = gam[c] * m.alpha_b[f] * fixed[f];
            dg_b += Db + dmax(Fbf, 0.0);
            rbn += (Db + dmax(-Fbf, 0.0)) * phib[f];
        }
        const double diag = ((dg_o + dg_n) + dg_b) + Sp[c];
        const double rhs = (rbn + rt) + Sc[c];
        diag_o[c] = diag;
        rhs_o[c] = rhs;
        const double dr = diag / relax;
        diag_r[c] = dr;
        inv_r[c] = 1.0 / dr;
        b[c] = rhs + crel * diag * pc;
        x0[c] = pc;
        mx = dmax(mx, diag);
        const bool keep = pc > floor_v * (1.0 + 1e-9) && !(held && held[c]);
        if (keep) {
            acc[0] += fabs(rhs - (diag * pc + ax));
            acc[1] += fabs(diag * pc);
        }
    }
    block_partials<2>(acc, sums);
    red[threadIdx.x] = mx;
    __syncthreads();
    for (int s = kBlock / 2; s > 0; s >>= 1) {
        if (threadIdx.x < s) red[threadIdx.x] = dmax(red[threadIdx.x], red[threadIdx.x + s]);
        __syncthreads();
    }
    if (threadIdx.x == 0) dmaxp[blockIdx.x] = red[0];
}

// out[0], out[1]: the residual sums; out[2]: max diag (one workgroup)
__global__ void __launch_bounds__(kBlock) k_sst_fin(const double* __restrict__ sums, const double* __restrict__ dmaxp,
                                                    int nparts, double* __restrict__ out) {
    __shared__ double red[kBlock];
    double tot[2];
    const bool t0 = finalize_sum<2>(sums, nparts, tot);
    double mx = -INFINITY;
    for (int i = threadIdx.x; i < nparts; i += kBlock) mx = dmax(mx, dmaxp[i]);
    red[threadIdx.x] = mx;
    __syncthreads();
    for (int s = kBlock / 2; s > 0; s >>= 1) {
        if (threadIdx.x < s) red[threadIdx.x] = dmax(red[threadIdx.x], red[threadIdx.x + s]);
        __syncthreads();
    }
    if (t0) {
        out[0] = tot[0];
        out[1] = tot[1];
        out[2] = red[0];
    }
}

// The wall cells' omega held at sqrt(omega_vis^2 + omega_log^2) by a penalty 1e10 max(diag)
// (turbulence.py:286-290, 390-397); k is the field before the update.
__global__ void __launch_bounds__(kBlock) k_sst_hold(int64_t nh, const int32_t* __restrict__ cells,
                                                     const double* __restrict__ ywc, const double* __restrict__ k,
                                                     const double* __restrict__ phi, const double* __restrict__ maxd,
                                                     double six_nu, double log_den0, double relax,
                                                     double* __restrict__ diag_r, const double* __restrict__ diag,
                                                     const double* __restrict__ rhs, double* __restrict__ b,
                                                     double* __restrict__ inv_r) {
    const int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    if (i >= nh) return;
    const int64_t c = cells[i];
    const double yc = ywc[i];
    const double vis = six_nu / (kB1 * (yc * yc));
    const double lg = sqrt(dmax(k[c], 0.0)) / (log_den0 * yc);
    const double val = sqrt(vis * vis + lg * lg);
    const double big = 1e10 * maxd[0];
    const double dgf = diag[c] + big;
    const double rf = rhs[c] + big * val;
    const double dr = dgf / relax;
    diag_r[c] = dr;
    inv_r[c] = 1.0 / dr;
    b[c] = rf + ((1.0 - relax) / relax) * dgf * phi[c];
}

// The new k and omega at their floors, then mu_t = rho a1 k / max(a1 omega, S F2) with F2 of
// the new fields (turbulence.py:302-303, 382-388; F2 needs no gradient).
__global__ void __launch_bounds__(kBlock) k_sst_update(int64_t n, const double* __restrict__ xk,
                                                       const double* __restrict__ xw, double kmin, double wmin,
                                                       const double* __restrict__ y, const double* __restrict__ S,
                                                       ftmi_sst_coef C, double* __restrict__ k, double* __restrict__ w,
                                                       double* __restrict__ mut) {
    int64_t lo, hi;
    block_range(n, lo, hi);
    for (int64_t c = lo + threadIdx.x; c < hi; c += kBlock) {
        const double kc = dmax(xk[c], kmin), wc = dmax(xw[c], wmin), yc = y[c];
        const double sk = sqrt(kc);
        const double arg2 = dmax(2.0 * sk / (kBstar * wc * yc), C.nu500 / (yc * yc * wc));
        const double F2 = tanh(arg2 * arg2);
        k[c] = kc;
        w[c] = wc;
        mut[c] = C.rho_a1 * kc / dmax(kA1 * wc, S[c] * F2);
    }
}

// Face diffusivities (turbulence.py:235-249): mode 0 momentum, c0 + mu_t on faces and c0 +
// mu_t(owner) on boundary faces (wall faces are then set by the wall function); mode 1 energy,
// c0 + mu_t / Pr_t, and c0 on wall faces.
__global__ void __launch_bounds__(kBlock) k_sst_face_gamma(int64_t nf, int64_t nb, const int32_t* __restrict__ P,
                                                           const int32_t* __restrict__ N,
                                                           const double* __restrict__ wgt,
                                                           const int32_t* __restrict__ Pb,
                                                           const int32_t* __restrict__ kind,
                                                           const double* __restrict__ mut, double c0, int mode,
                                                           double* __restrict__ gf, double* __restrict__ gb) {
    int64_t lo, hi;
    block_range(nf + nb, lo, hi);
    for (int64_t i = lo + threadIdx.x; i < hi; i += kBlock) {
        if (i < nf) {
            const double mf = wgt[i] * mut[P[i]] + (1.0 - wgt[i]) * mut[N[i]];
            gf[i] = mode == 0 ? c0 + mf : c0 + mf / kPrT;
        } else {
            const int64_t f = i - nf;
            const double mb = mut[Pb[f]];
            if (mode == 0)
                gb[f] = c0 + mb;
            else
                gb[f] = c0 + (kind[f] == WALL ? 0.0 : mb / kPrT);
        }
    }
}

__global__ void k_zero_int(int* x) { *x = 0; }

// Spalding's law for every boundary face (turbulence.py:65-87, 399-417). The reference iterates
// Newton on all faces together until every step is small; here pass 1 finds, per face, the first
// iteration whose step meets the test and takes the maximum over faces, and pass 2 runs every
// face exactly that many iterations, which is the same result wherever a face's test, once met,
// stays met (Newton's steps shrink).

struct Spalding {
    double eb;  // exp(-kappa B)
    __device__ double start(double r) const {
        return r < 120.0 ? sqrt(dmax(r, 1e-300)) : log(dmax(kEWall * sqrt(dmax(r, 1.0)), 1.0)) / kKappa + 1.0;
    }
    // one Newton step: returns the step, updates u
    __device__ double step(double r, double& u) const {
        const double ku = dmin(kKappa * u, 60.0);
        const double ex = exp(ku);
        const double g = u + eb * (ex - 1 - ku - ku * ku / 2 - pow(ku, 3.0) / 6);
        const double dg = 1 + eb * kKappa * (ex - 1 - ku - ku * ku / 2);
        const double uu = dmax(u, 1e-300);
        const double f = r / uu - g;
        const double df = -r / (uu * uu) - dg;
        const double st = f / df;
        u = dmax(u - st, 0.5 * u);
        return st;
    }
};

__device__ inline double wall_re_y(int64_t f, int64_t nb, int dim, const int32_t* Pb, const double* U, int64_t nc,
                                   const double* ubf, const double* nrm, const double* yb, double nu) {
    const int64_t c = Pb[f];
    double rel[3], rn = 0.0;
    for (int a = 0; a < dim; ++a) rel[a] = U[a * nc + c] - ubf[f * dim + a];
    rn = rel[0] * nrm[f * dim];
    for (int a = 1; a < dim; ++a) rn += rel[a] * nrm[f * dim + a];
    double t0 = rel[0] - rn * nrm[f * dim];
    double s = t0 * t0;
    for (int a = 1; a < dim; ++a) {
        const double t = rel[a] - rn * nrm[f * dim + a];
        s += t * t;
    }
    return sqrt(s) * yb[f] / nu;
}

__global__ void __launch_bounds__(kBlock) k_spalding_count(int64_t nb, int dim, const int32_t* __restrict__ Pb,
                                                           const double* __restrict__ U, int64_t nc,
                                                           const double* __restrict__ ubf,
                                                           const double* __restrict__ nrm,
                                                           const double* __restrict__ yb, double nu, Spalding sp,
                                                           int* __restrict__ iters) {
    const int64_t f = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    if (f >= nb) return;
    const double r = dmax(wall_re_y(f, nb, dim, Pb, U, nc, ubf, nrm, yb, nu), 0.0);
    double u = sp.start(r);
    int it = 30;
    for (int i = 1; i <= 30; ++i) {
        const double st = sp.step(r, u);
        if (fabs(st) <= 1e-10 * dmax(u, 1e-12)) {
            it = i;
            break;
        }
    }
    atomicMax(iters, it);
}

__global__ void __launch_bounds__(kBlock) k_spalding_apply(int64_t nb, int dim, const int32_t* __restrict__ Pb,
                                                           const int32_t* __restrict__ kind,
                                                           const double* __restrict__ U, int64_t nc,
                                                           const double* __restrict__ ubf,
                                                           const double* __restrict__ nrm,
                                                           const double* __restrict__ yb, double nu, double mu,
                                                           Spalding sp, const int* __restrict__ iters,
                                                           double* __restrict__ gb, double* __restrict__ yplus) {
    const int64_t f = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    if (f >= nb) return;
    const double r = dmax(wall_re_y(f, nb, dim, Pb, U, nc, ubf, nrm, yb, nu), 0.0);
    double u = sp.start(r);
    const int n_it = *iters;
    for (int i = 0; i < n_it; ++i) sp.step(r, u);
    const double up = r > 0.0 ? u : 0.0;
    const bool ok = r > 1e-12 && up > 0.0;
    const bool wall = kind[f] == WALL;
    if (wall) gb[f] = ok ? mu * r / (up * up) : mu;
    yplus[f] = (wall && ok) ? r / up : 0.0;
}

// Transpose stress div(mu_t (grad U)^T) per cell (turbulence.py:251-261); out (DIM, nc).
template <int DIM>
__global__ void __launch_bounds__(kBlock) k_sst_tstress(ftmi_mesh m, const double* __restrict__ mut,
                                                        const double* __restrict__ gU, double* __restrict__ out) {
    int64_t lo, hi;
    block_range(m.nc, lo, hi);
    const int64_t n = m.nc;
    for (int64_t c = lo + threadIdx.x; c < hi; c += kBlock) {
        double acc[DIM];
#pragma unroll
        for (int q = 0; q < DIM; ++q) acc[q] = 0.0;
        const int32_t e = m.cptr[c + 1];
        for (int32_t h = m.cptr[c]; h < e; ++h) {
            const int32_t code = m.cface[h];
            const int64_t f = code >> 1;
            const bool isN = code & 1;
            const int64_t o = m.cnbr[h];
            const int64_t P = isN ? o : c, N = isN ? c : o;
            const double wf = m.w[f], wn = 1.0 - wf;
            const double mf = wf * mut[P] + wn * mut[N];
#pragma unroll
            for (int q = 0; q < DIM; ++q) {
                // column q of grad U: d u_j / d x_q, j = 0..DIM-1, interpolated, dotted with S
                double s = (wf * gU[(0 * n + P) * DIM + q] + wn * gU[(0 * n + N) * DIM + q]) * m.S[f * DIM];
#pragma unroll
                for (int j = 1; j < DIM; ++j)
                    s += (wf * gU[(j * n + P) * DIM + q] + wn * gU[(j * n + N) * DIM + q]) * m.S[f * DIM + j];
                const double v = mf * s;
                acc[q] += isN ? -v : v;
            }
        }
#pragma unroll
        for (int q = 0; q < DIM; ++q) out[q * n + c] = acc[q];
    }
}

}  // namespace

extern "C" {

int ftmi_sst_bvals(const ftmi_mesh* mesh, const ftmi_bc* bc, const double* Fb, const double* k, const double* w,
                   const double* k_in, const double* w_in, double* kb, double* wb, double* fixed, void* stream) {
    if (mesh->nb > 0)
        k_sst_bvals<<<grid_for(mesh->nb), kBlock, 0, (hipStream_t)stream>>>(mesh->nb, mesh->Pb, bc->kind, Fb, k, w,
                                                                             k_in, w_in, kb, wb, fixed);
    FTMI_LAUNCH_CHECK("ftmi_sst_bvals");
    return 0;
}

int ftmi_sst_cell(const ftmi_mesh* mesh, const ftmi_sst_coef* coef, const double* gU, const double* gk,
                  const double* gw, const double* k, const double* w, const double* mut, const double* y,
                  double* S, double* gam_k, double* sck, double* spk, double* gam_w, double* scw, double* spw,
                  void* stream) {
    const int g = grid_for(mesh->nc);
    hipStream_t s = (hipStream_t)stream;
    if (mesh->dim == 2)
        k_sst_cell<2><<<g, kBlock, 0, s>>>(mesh->nc, gU, gk, gw, k, w, mut, y, mesh->vol, *coef, S, gam_k, sck, spk,
                                          gam_w, scw, spw);
    else if (mesh->dim == 3)
        k_sst_cell<3><<<g, kBlock, 0, s>>>(mesh->nc, gU, gk, gw, k, w, mut, y, mesh->vol, *coef, S, gam_k, sck, spk,
                                          gam_w, scw, spw);
    else {
        set_error_msg("ftmi_sst_cell: dim must be 2 or 3");
        return -1;
    }
    FTMI_LAUNCH_CHECK("ftmi_sst_cell");
    return 0;
}

int ftmi_sst_work_doubles(void) { return 3 * kMaxGrid + 3; }

int ftmi_sst_assemble(const ftmi_mesh* mesh, const double* F, const double* Fb, const double* gam, const double* phi,
                      const double* grad, const double* phib, const double* fixed, const double* Sc, const double* Sp,
                      const uint8_t* held, double floor_v, double relax, double* diag, double* rhs, double* diag_r,
                      double* off, double* b, double* inv_r, double* x0, double* work, double* out3, void* stream) {
    const int g = grid_part(mesh->nc);
    hipStream_t s = (hipStream_t)stream;
    double* sums = work;
    double* dmaxp = work + 2 * kMaxGrid;
    if (mesh->dim == 2)
        k_sst_assemble<2><<<g, kBlock, 0, s>>>(*mesh, F, Fb, gam, phi, grad, phib, fixed, Sc, Sp, held, floor_v,
                                              relax, diag, rhs, diag_r, off, b, inv_r, x0, sums, dmaxp);
    else if (mesh->dim == 3)
        k_sst_assemble<3><<<g, kBlock, 0, s>>>(*mesh, F, Fb, gam, phi, grad, phib, fixed, Sc, Sp, held, floor_v,
                                              relax, diag, rhs, diag_r, off, b, inv_r, x0, sums, dmaxp);
    else {
        set_error_msg("ftmi_sst_assemble: dim must be 2 or 3");
        return -1;
    }
    k_sst_fin<<<1, kBlock, 0, s>>>(sums, dmaxp, g, out3);
    FTMI_LAUNCH_CHECK("ftmi_sst_assemble");
    return 0;
}

int ftmi_sst_hold(int64_t nh, const int32_t* cells, const double* y_wall_cell, const double* k, const double* phi,
                  const double* out3, double nu, double relax, const double* diag, const double* rhs, double* diag_r,
                  double* b, double* inv_r, void* stream) {
    if (nh <= 0) return 0;
    // the constants of omega_vis = 6 nu / (beta_1 y^2) and omega_log = sqrt(k) / (C_mu^1/4 kappa y)
    const double log_den0 = std::pow(0.09, 0.25) * kKappa;
    k_sst_hold<<<(int)((nh + kBlock - 1) / kBlock), kBlock, 0, (hipStream_t)stream>>>(
        nh, cells, y_wall_cell, k, phi, out3 + 2, 6.0 * nu, log_den0, relax, diag_r, diag, rhs, b, inv_r);
    FTMI_LAUNCH_CHECK("ftmi_sst_hold");
    return 0;
}

int ftmi_sst_update(int64_t n, const double* xk, const double* xw, double kmin, double wmin, const double* y,
                    const double* S, const ftmi_sst_coef* coef, double* k, double* w, double* mut, void* stream) {
    if (n > 0)
        k_sst_update<<<grid_for(n), kBlock, 0, (hipStream_t)stream>>>(n, xk, xw, kmin, wmin, y, S, *coef, k, w, mut);
    FTMI_LAUNCH_CHECK("ftmi_sst_update");
    return 0;
}

int ftmi_sst_face_gamma(const ftmi_mesh* mesh, const ftmi_bc* bc, const double* mut, double c0, int mode,
                        double* gam_f, double* gam_b, void* stream) {
    const int64_t tot = mesh->nf + mesh->nb;
    if (tot > 0)
        k_sst_face_gamma<<<grid_for(tot), kBlock, 0, (hipStream_t)stream>>>(
            mesh->nf, mesh->nb, mesh->P, mesh->N, mesh->w, mesh->Pb, bc->kind, mut, c0, mode, gam_f, gam_b);
    FTMI_LAUNCH_CHECK("ftmi_sst_face_gamma");
    return 0;
}

int ftmi_sst_wall(const ftmi_mesh* mesh, const ftmi_bc* bc, const double* U, const double* yb, double mu, double nu,
                  int* iters, double* gam_b, double* yplus, void* stream) {
    const int64_t nb = mesh->nb;
    if (nb <= 0) return 0;
    hipStream_t s = (hipStream_t)stream;
    Spalding sp{std::exp(-kKappa * 5.2)};
    const int g = (int)((nb + kBlock - 1) / kBlock);
    k_zero_int<<<1, 1, 0, s>>>(iters);  // a kernel, not a memset node, so the call can be captured
    k_spalding_count<<<g, kBlock, 0, s>>>(nb, (int)mesh->dim, mesh->Pb, U, mesh->nc, bc->ub_fixed, mesh->nb_unit, yb,
                                          nu, sp, iters);
    k_spalding_apply<<<g, kBlock, 0, s>>>(nb, (int)mesh->dim, mesh->Pb, bc->kind, U, mesh->nc, bc->ub_fixed,
                                          mesh->nb_unit, yb, nu, mu, sp, iters, gam_b, yplus);
    FTMI_LAUNCH_CHECK("ftmi_sst_wall");
    return 0;
}

int ftmi_sst_tstress(const ftmi_mesh* mesh, const double* mut, const double* gU, double* out, void* stream) {
    const int g = grid_for(mesh->nc);
    hipStream_t s = (hipStream_t)stream;
    if (mesh->dim == 2)
        k_sst_tstress<2><<<g, kBlock, 0, s>>>(*mesh, mut, gU, out);
    else if (mesh->dim == 3)
        k_sst_tstress<3><<<g, kBlock, 0, s>>>(*mesh, mut, gU, out);
    else {
        set_error_msg("ftmi_sst_tstress: dim must be 2 or 3");
        return -1;
    }
    FTMI_LAUNCH_CHECK("ftmi_sst_tstress");
    return 0;
}

}  // extern "C"
