// Device-side numerics of the explicit update: primitive recovery, ghost
// cells, flux splitting, 5th-order reconstructions and the characteristic
// projection, in FP64.  The formulas are the reference's; divisions in the
// per-face and per-cell paths go through refined reciprocals and the library
// is built with FMA contraction, so results agree with the reference's to
// rounding level rather than bit for bit.  Citations are relative to the
// reference repository root.
#pragma once
#include "dense_internal.h"

namespace dense {

// Single-precision literals of the reference, widened as Fortran widens them.
constexpr double P_FLOOR = (double)1.0e-6f;        // get_primary.f90:58
constexpr double WENO_SS = (double)1.0e-6f;        // weno_mod.f90:37
constexpr double TWO_THIRDS_SP = (double)(2.0f / 3.0f);  // time_advance_general.f90:242

// weno_mod.f90:3-14
constexpr double A300 = 1.0 / 3.0, A301 = -7.0 / 6.0, A302 = 11.0 / 6.0;
constexpr double A310 = -1.0 / 6.0, A311 = 5.0 / 6.0, A312 = 1.0 / 3.0;
constexpr double A320 = 1.0 / 3.0, A321 = 5.0 / 6.0, A322 = -1.0 / 6.0;
constexpr double C30 = 0.1, C31 = 0.6, C32 = 0.3;

enum Recon : int { R_UP5 = 0, R_JS = 1, R_Z = 2 };

// ---------------------------------------------------------------- division
// gfx950 has no f64 divide: an IEEE a/b expands to 11 instructions around a
// quarter-rate v_rcp_f64 (div_scale twice, two Newton steps, div_fmas,
// div_fixup).  The hot paths divide through reciprocals instead, and the
// library is compiled with FMA contraction: results agree with the
// reference's step-by-step IEEE evaluation to rounding level (the tolerance
// the workloads state is 1e-11 relative on smooth flows and 1e-9 density L1
// on shocked flows), not bit for bit.
//   rcp(b): v_rcp_f64 refined by two Newton steps (within 1 ulp of 1/b for
//           normal b whose reciprocal is normal).
//   fdiv(a, b): a*rcp(b) corrected by one residual step (the quotient to
//           within an ulp for normal operands away from the exponent limits).
__device__ __forceinline__ double rcp(double b)
{
    const double r = __builtin_amdgcn_rcp(b);
    const double e0 = fma(-b, r, 1.0);
    const double y1 = fma(r, e0, r);
    const double e1 = fma(-b, y1, 1.0);
    return fma(y1, e1, y1);
}

__device__ __forceinline__ double fdiv(double a, double b)
{
    const double y = rcp(b);
    const double q0 = a * y;
    return fma(fma(-b, q0, a), y, q0);
}

// sqrt(x) and 1/sqrt(x) of a positive normal x from v_rsq_f64, refined by the
// coupled Newton iteration for (sqrt(x), 1/(2 sqrt(x))) and a last residual
// step on sqrt(x): within an ulp, without the generic square root's scaling
// and special-case code (x here is a density or a squared sound speed).
__device__ __forceinline__ void sqrt_rsq(double x, double& sq, double& rs)
{
    const double r = __builtin_amdgcn_rsq(x);
    double g = x * r, h = 0.5 * r;
    const double e = fma(-g, h, 0.5);
    g = fma(g, e, g);
    h = fma(h, e, h);
    const double d = fma(-g, g, x);
    sq = fma(d, h, g);
    rs = 2.0 * h;
}

// ---------------------------------------------------------------- stencils

// smoothness indicators, weno_mod.f90:152-159 / :175-182:
// beta = 13/12 x^2 + 1/4 y^2
__device__ __forceinline__ void smooth_ind(const double* lf, double& is0, double& is1, double& is2)
{
    constexpr double C13 = 13.0 / 12.0;
    double x, y;
    x = (lf[0] - 2.0 * lf[1]) + lf[2];
    y = (lf[0] - 4.0 * lf[1]) + 3.0 * lf[2];
    is0 = C13 * (x * x) + 0.25 * (y * y);
    x = (lf[1] - 2.0 * lf[2]) + lf[3];
    y = lf[1] - lf[3];
    is1 = C13 * (x * x) + 0.25 * (y * y);
    x = (lf[2] - 2.0 * lf[3]) + lf[4];
    y = (3.0 * lf[2] - 4.0 * lf[3]) + lf[4];
    is2 = C13 * (x * x) + 0.25 * (y * y);
}

// candidate polynomials, weno_mod.f90:57-59
__device__ __forceinline__ void candidates(const double* lf, double& q0, double& q1, double& q2)
{
    q0 = A300 * lf[0] + A301 * lf[1] + A302 * lf[2];
    q1 = A310 * lf[1] + A311 * lf[2] + A312 * lf[3];
    q2 = A320 * lf[2] + A321 * lf[3] + A322 * lf[4];
}

// WENO-Z unnormalised weights a_k = c_k (1 + (tau5/(eps + beta_k))^2),
// tau5 = |beta_2 - beta_0| (weno_mod.f90:184-188), for the indicator and
// theta of AdaWENO (which need the sum of these weights itself).
__device__ __forceinline__ void weno5z_alpha(const double* lf, double& aa0, double& aa1, double& aa2)
{
    double is0, is1, is2;
    smooth_ind(lf, is0, is1, is2);
    const double tau5 = fabs(is2 - is0);
    const double t0 = tau5 * rcp(WENO_SS + is0);
    const double t1 = tau5 * rcp(WENO_SS + is1);
    const double t2 = tau5 * rcp(WENO_SS + is2);
    aa0 = C30 + C30 * (t0 * t0);
    aa1 = C31 + C31 * (t1 * t1);
    aa2 = C32 + C32 * (t2 * t2);
}

// WENO-Z normalised weights w_k = a_k / sum a and AdaWENO's indicator
// theta = 1/(1 + (sum a - 1)^2) (weno_mod.f90:67-95)
__device__ __forceinline__ void weno5z_omega(const double* lf, double& w0, double& w1, double& w2, double* theta)
{
    double aa0, aa1, aa2;
    weno5z_alpha(lf, aa0, aa1, aa2);
    const double s = aa0 + aa1 + aa2;
    const double ys = rcp(s);
    w0 = aa0 * ys;
    w1 = aa1 * ys;
    w2 = aa2 * ys;
    if (theta) {
        const double d = s - 1.0;
        *theta = rcp(1.0 + d * d);
    }
}

// weno5z_stencil, weno_mod.f90:50-65: sum w_k q_k with w_k = a_k / sum a_k.
// The weights are scale-free, so every a_k is multiplied by (d0 d1 d2)^2,
// d_k = eps + beta_k:  a_k (d0 d1 d2)^2 = c_k (d_k^2 + tau5^2) e_k^2 with
// e_k = d0 d1 d2 / d_k the product of the other two -- no quotient left but
// the one normalisation.  d_k >= eps = 1e-6 and beta_k is a squared flux
// difference, so the products stay far inside the FP64 range.
__device__ __forceinline__ double weno5z(const double* lf)
{
    double q0, q1, q2, is0, is1, is2;
    candidates(lf, q0, q1, q2);
    smooth_ind(lf, is0, is1, is2);
    const double tau5 = fabs(is2 - is0), t2 = tau5 * tau5;
    const double d0 = WENO_SS + is0, d1 = WENO_SS + is1, d2 = WENO_SS + is2;
    const double e0 = d1 * d2, e1 = d0 * d2, e2 = d0 * d1;
    const double aa0 = C30 * (d0 * d0 + t2) * (e0 * e0);
    const double aa1 = C31 * (d1 * d1 + t2) * (e1 * e1);
    const double aa2 = C32 * (d2 * d2 + t2) * (e2 * e2);
    return (aa0 * q0 + aa1 * q1 + aa2 * q2) * rcp(aa0 + aa1 + aa2);
}

// weno5js_stencil, weno_mod.f90:98-165: a_k = c_k/(eps+beta_k)^2, scaled by
// (d0 d1 d2)^2 as above: c_k e_k^2, one reciprocal per stencil.
__device__ __forceinline__ double weno5js(const double* lf)
{
    double q0, q1, q2, is0, is1, is2;
    candidates(lf, q0, q1, q2);
    smooth_ind(lf, is0, is1, is2);
    const double d0 = WENO_SS + is0, d1 = WENO_SS + is1, d2 = WENO_SS + is2;
    const double e0 = d1 * d2, e1 = d0 * d2, e2 = d0 * d1;
    const double aa0 = C30 * (e0 * e0), aa1 = C31 * (e1 * e1), aa2 = C32 * (e2 * e2);
    return (aa0 * q0 + aa1 * q1 + aa2 * q2) * rcp(aa0 + aa1 + aa2);
}

// scheme_upwind5_general.f90:86-90: (2, -13, 47, 27, -3) / 60
__device__ __forceinline__ double upwind5(const double* lf)
{
    constexpr double R60 = 1.0 / 60.0;
    return (2.0 * lf[0] - 13.0 * lf[1] + 47.0 * lf[2] + 27.0 * lf[3] - 3.0 * lf[4]) * R60;
}

template <int R>
__device__ __forceinline__ double recon(const double* lf)
{
    if constexpr (R == R_UP5) return upwind5(lf);
    else if constexpr (R == R_JS) return weno5js(lf);
    else return weno5z(lf);
}

// ---------------------------------------------------------------- cells

// Primitive variables of one cell, get_primary.f90:52-61.  pr = (u, v, p, c).
// WITH_C = false leaves out the sound speed (pr[3] = 0) for callers that
// never read it.
template <bool WITH_C = true>
__device__ __forceinline__ void prim_of(double gamma, const double* c, double* pr)
{
    const double rho = c[0];
    const double ir = rcp(rho);
    const double u = c[1] * ir;
    const double v = c[2] * ir;
    double p = (gamma - 1.0) * (c[3] - 0.5 * rho * (u * u + v * v));
    if (p < 0.0) p = P_FLOOR;
    pr[0] = u;
    pr[1] = v;
    pr[2] = p;
    pr[3] = WITH_C ? sqrt(gamma * p * ir) : 0.0;
}

// Dirichlet ghost from a ghost record (conserved state, then u, v, p, c; bc_fill)
__device__ __forceinline__ void from_record(const double* r, double* c, double* pr)
{
#pragma unroll
    for (int v = 0; v < 4; v++) {
        c[v] = r[v];
        pr[v] = r[4 + v];
    }
}

template <bool WITH_C = true>
__device__ __forceinline__ void load_interior(const StageArgs& a, const double* __restrict__ U, int gx, int gy,
                                              double* c, double* pr)
{
    const long long o = (long long)gy * a.ld + gx;
#pragma unroll
    for (int k = 0; k < 4; k++) c[k] = U[k * a.plane + o];
    prim_of<WITH_C>(a.gamma, c, pr);
}

// Ghost cell (gx,gy) outside the interior in exactly one direction, k layers
// deep (k = 1 adjacent).  Reproduces the reference's boundary routines: the
// ghost conserved state AND the ghost primitive state, which the reference
// sets independently (e.g. extrapolated separately, boundary_ivc_general.f90:72-80).
__device__ __forceinline__ void ghost_cell(const StageArgs& a, const double* __restrict__ U, int gx, int gy, double t,
                                        double* c, double* pr)
{
    const int ni = a.ni, nj = a.nj;
    const bool xside = (gx < 0 || gx >= ni);
    int side, k;
    if (xside) {
        side = gx < 0 ? DENSE_SIDE_LEFT : DENSE_SIDE_RIGHT;
        k = gx < 0 ? -gx : gx - ni + 1;
    } else {
        side = gy < 0 ? DENSE_SIDE_BOTTOM : DENSE_SIDE_TOP;
        k = gy < 0 ? -gy : gy - nj + 1;
    }
    const bool low = (side == DENSE_SIDE_LEFT || side == DENSE_SIDE_BOTTOM);
    const int n = xside ? ni : nj;
    int kind = a.bc.kind[side];

    if (kind == DENSE_BC_DMR_WALL) {
        // x of the cell column, x(index_i-bfsize_i,1) = (i - 0.5d0)*dx
        const double x = ((double)(gx + 1) - 0.5) * a.dx;
        kind = (x >= 1.0 / 6.0) ? DENSE_BC_REFLECT : DENSE_BC_DIRICHLET;
    }
    switch (kind) {
    case DENSE_BC_PERIODIC: {
        const int src = low ? n - k : k - 1;
        if (xside) load_interior(a, U, src, gy, c, pr);
        else load_interior(a, U, gx, src, c, pr);
        return;
    }
    case DENSE_BC_REFLECT: {
        const int src = low ? k - 1 : n - k;
        if (xside) {
            load_interior(a, U, src, gy, c, pr);
            c[1] = -c[1];
            pr[0] = -pr[0];
        } else {
            load_interior(a, U, gx, src, c, pr);
            c[2] = -c[2];
            pr[1] = -pr[1];
        }
        return;
    }
    case DENSE_BC_EXTRAP: {
        // ghost(k) = ghost(k-1)*4/3 - ghost(k-2)/3, starting from the two
        // boundary cells (boundary_ivc_general.f90:64-67, loop :63 outward)
        const int i0 = low ? 0 : n - 1, i1 = low ? 1 : n - 2;
        double q1[8], q2[8];
        if (xside) {
            load_interior(a, U, i0, gy, q1, q1 + 4);
            load_interior(a, U, i1, gy, q2, q2 + 4);
        } else {
            load_interior(a, U, gx, i0, q1, q1 + 4);
            load_interior(a, U, gx, i1, q2, q2 + 4);
        }
        for (int s = 0; s < k; s++) {
#pragma unroll
            for (int v = 0; v < 8; v++) {
                const double g = q1[v] * 4.0 / 3.0 - q2[v] / 3.0;
                q2[v] = q1[v];
                q1[v] = g;
            }
        }
#pragma unroll
        for (int v = 0; v < 4; v++) {
            c[v] = q1[v];
            pr[v] = q1[4 + v];
        }
        return;
    }
    case DENSE_BC_SSL_INLET: {
        // boundary_ssl_general.f90:55-85 at t^n: rho, p by the sign of y;
        // u = 2.5 + 0.5 tanh(2y) and exp(-y^2/10) tabulated per row by the host,
        // 0.05 cos(.) of the two forcing modes in st->bc_t (host libm)
        // (p/(gamma-1) and the two sound speeds from bc_fill)
        const double y = ((double)(gy + 1) - 0.5) * a.dy - a.bc.state[side][0];
        const double e = a.bc_rows[a.nj + gy];
        const double v = a.st->bc_t[0] * e + a.st->bc_t[1] * e;
        const double u = a.bc_rows[gy];
        const double rho = y >= 0.0 ? SSL_RHO_HI : SSL_RHO_LO;
        c[0] = rho;
        c[1] = rho * u;
        c[2] = rho * v;
        c[3] = a.bc.ssl_e + 0.5 * rho * (u * u + v * v);
        pr[0] = u;
        pr[1] = v;
        pr[2] = SSL_P;
        pr[3] = y >= 0.0 ? a.bc.ssl_c[0] : a.bc.ssl_c[1];
        return;
    }
    case DENSE_BC_DMR_TOP: {
        // boundary_doublemach_general.f90:67-80, shock at t^n
        const double x = ((double)(gx + 1) - 0.5) * a.dx;
        from_record(x < (1.0 / 6.0 + (1.0 + 20.0 * t) / a.sqrt3) ? a.bc.rec[side] : a.bc.pre, c, pr);
        return;
    }
    default:  // DIRICHLET
        from_record(a.bc.rec[side], c, pr);
        return;
    }
}

// ---------------------------------------------------------------- fluxes

// Lax-Friedrichs split of one cell along DIR, sign +1/-1, doubled:
// 2F = (u_n +/- alpha)*U + (0, p e_n, u_n p), where the reference forms
// F = 0.5*(...) (flux_splitting_LLF.f90:39-47 (x), :64-72 (y); GLF
// flux_splitting_GLF.f90:31-43).  Multiplying by 0.5 is exact and commutes with
// rounding, also inside an FMA, so the caller folds the 0.5 into the linear map
// that consumes 2F and gets the reference's values bit for bit.
template <int DIR>
__device__ __forceinline__ void lf_flux2(double upa, double un, double p, const double* c, double* f)
{
    constexpr int NRM = DIR == 0 ? 1 : 2;
    f[0] = upa * c[0];
    f[NRM] = upa * c[NRM] + p;
    f[3 - NRM] = upa * c[3 - NRM];
    f[3] = upa * c[3] + un * p;
}

// Steger-Warming split of one cell along DIR, both signs,
// flux_splitting_SW.f90:36-92 (x), :103-161 (y).
template <int DIR>
__device__ __forceinline__ void sw_flux(double gamma, double rho, double u, double v, double alpha, double* fp,
                                        double* fm)
{
    const double epsl = 0.125;
    const double un = DIR == 0 ? u : v;
    const double uu2 = u * u + v * v;
    const double lam[3] = {un, un - alpha, un + alpha};
    const double g2 = 2.0 * (gamma - 1.0);
    const double r2g = rho * rcp(2.0 * gamma);        // rho / (2 gamma)
    const double wc = (3.0 - gamma) * (alpha * alpha) * rcp(g2);  // w term factor, :51
    double sq[3];
#pragma unroll
    for (int q = 0; q < 3; q++) sq[q] = sqrt(lam[q] * lam[q] + epsl * epsl);
#pragma unroll
    for (int sgn = 0; sgn < 2; sgn++) {
        double* f = sgn == 0 ? fp : fm;
        double lp[3];
#pragma unroll
        for (int q = 0; q < 3; q++) lp[q] = sgn == 0 ? 0.5 * (lam[q] + sq[q]) : 0.5 * (lam[q] - sq[q]);
        const double wii = (lp[1] + lp[2]) * wc;
        f[0] = (lp[0] * g2 + lp[1] + lp[2]) * r2g;
        if (DIR == 0) {
            f[1] = (lp[0] * g2 * u + lp[1] * (u - alpha) + lp[2] * (u + alpha)) * r2g;
            f[2] = (lp[0] * g2 * v + lp[1] * v + lp[2] * v) * r2g;
            f[3] = (lp[0] * (gamma - 1.0) * uu2 + lp[1] * 0.5 * ((u - alpha) * (u - alpha) + v * v) +
                    lp[2] * 0.5 * ((u + alpha) * (u + alpha) + v * v) + wii) *
                   r2g;
        } else {
            f[1] = (lp[0] * g2 * u + lp[1] * u + lp[2] * u) * r2g;
            f[2] = (lp[0] * g2 * v + lp[1] * (v - alpha) + lp[2] * (v + alpha)) * r2g;
            f[3] = (lp[0] * (gamma - 1.0) * uu2 + lp[1] * 0.5 * ((v - alpha) * (v - alpha) + u * u) +
                    lp[2] * 0.5 * ((v + alpha) * (v + alpha) + u * u) + wii) *
                   r2g;
        }
    }
}

}  // namespace dense
