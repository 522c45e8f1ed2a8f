// Device-side numerics of the explicit update: primitive recovery, ghost
// cells, flux splitting, 5th-order reconstructions and the characteristic
// projection.  Every expression keeps the reference's evaluation order and
// the library is built with -ffp-contract=off, so results are bitwise equal
// to the reference's FP64 arithmetic.  Citations are relative to the
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
// gfx950 has no f64 divide; the compiler expands a/b into
//   s0 = div_scale(b, b, a); y = rcp(s0) refined by two Newton steps
//   (e = fma(-s0, y, 1), y = fma(y, e, y), twice); s1 = div_scale(a, b, a);
//   q0 = s1*y; r = fma(-s0, q0, s1); q = div_fmas(r, y, q0); div_fixup(q, b, a).
// Both div_scale steps return their operand unscaled and leave the scale flag
// clear, and div_fixup returns its input, when the operands and the quotient
// are normal numbers away from the exponent limits:
//   2^-969 <= |a|, b normal with |b| < 2^1021, 2^-1022 <= |a/b| < 2^767,
// or a = +-0 with b > 0 (div_fixup returns a zero of a's sign).  In that range
// a/b is, bit for bit, y = nr_recip(b); q = div_by(a, b, y), so a divisor
// shared by several quotients needs one reciprocal.  Every use below tests a
// sufficient range condition per quotient and recomputes that quotient with
// '/' where it fails (a short divergent branch, taken on extreme data only).
__device__ __forceinline__ double nr_recip(double b)
{
    const double r = __builtin_amdgcn_rcp(b);
    const double e0 = fma(-b, r, 1.0);
    const double y1 = fma(r, e0, r);
    const double e1 = fma(-b, y1, 1.0);
    return fma(y1, e1, y1);
}

// The residual is formed negated, fma(b, q0, -a) = -fma(-b, q0, a) exactly,
// so that a = -0 yields -0 like IEEE division (identical bits otherwise).
__device__ __forceinline__ double div_by(double a, double b, double y)
{
    const double q0 = a * y;
    const double rn = fma(b, q0, -a);
    return fma(-rn, y, q0);
}

// n/12 for n >= 0 (the smoothness indicators).  The residual is formed at
// scale 2^128: for 2^-969 <= n <= 2^890 this is the sequence above with an
// exact residual scaled by a power of two (identical bits); for tiny n
// (2^-1018 <= n < 2^-969, where the compiler's sequence scales both operands
// by 2^128) it is that scaled sequence, q = RN(n*2^128/12)*2^-128 = RN(n/12),
// the quotient being normal.  n = 0 gives +0.  Other n take '/'.
// 12 is hidden from constant folding so that its reciprocal is formed at run
// time by the same instructions as in the compiler's n/12.0.
__device__ __forceinline__ double div12(double n)
{
    double twelve = 12.0;
    asm("" : "+v"(twelve));
    const double y = nr_recip(twelve);
    const double q0 = n * y;
    const double rn = fma(twelve * 0x1p128, q0, -(n * 0x1p128));
    double q = fma(-rn, y * 0x1p-128, q0);
    if (__builtin_expect(!(n == 0.0 || (n >= 0x1p-1018 && n <= 0x1p890)), 0)) q = n / 12.0;
    return q;
}

// ---------------------------------------------------------------- stencils

// The stencils take FAST = true in the characteristic WENO-Z scheme (divisions
// are about half of its instructions; measured 3-4 % faster per stage) and
// FAST = false ('/') elsewhere: the component-wise schemes run at the register
// limit of their occupancy, and AdaWENO measured no gain.
template <bool FAST>
__device__ __forceinline__ double over12(double n)
{
    if constexpr (FAST) return div12(n);
    else return n / 12.0;
}

// smoothness indicators, weno_mod.f90:152-159 / :175-182
template <bool FAST>
__device__ __forceinline__ void smooth_ind(const double* lf, double& is0, double& is1, double& is2)
{
    double x, y;
    x = (lf[0] - 2.0 * lf[1]) + lf[2];
    y = (lf[0] - 4.0 * lf[1]) + 3.0 * lf[2];
    is0 = over12<FAST>(13.0 * (x * x)) + (y * y) / 4.0;
    x = (lf[1] - 2.0 * lf[2]) + lf[3];
    y = lf[1] - lf[3];
    is1 = over12<FAST>(13.0 * (x * x)) + (y * y) / 4.0;
    x = (lf[2] - 2.0 * lf[3]) + lf[4];
    y = (3.0 * lf[2] - 4.0 * lf[3]) + lf[4];
    is2 = over12<FAST>(13.0 * (x * x)) + (y * y) / 4.0;
}

// candidate polynomials, weno_mod.f90:57-59
__device__ __forceinline__ void candidates(const double* lf, double& q0, double& q1, double& q2)
{
    q0 = A300 * lf[0] + A301 * lf[1] + A302 * lf[2];
    q1 = A310 * lf[1] + A311 * lf[2] + A312 * lf[3];
    q2 = A320 * lf[2] + A321 * lf[3] + A322 * lf[4];
}

// w_k = a_k / s for a_k = c_k(...) >= c_k >= 0.1 and s = sum a_k: one shared
// reciprocal when s <= 2^1000 (then 0.1 <= a_k <= s and the quotients lie in
// [2^-1004, 1]); '/' otherwise (s infinite or NaN).
template <bool FAST>
__device__ __forceinline__ void normalize3(double aa0, double aa1, double aa2, double s, double& w0, double& w1,
                                           double& w2)
{
    if constexpr (!FAST) {
        w0 = aa0 / s;
        w1 = aa1 / s;
        w2 = aa2 / s;
        return;
    }
    const double ys = nr_recip(s);
    w0 = div_by(aa0, s, ys);
    w1 = div_by(aa1, s, ys);
    w2 = div_by(aa2, s, ys);
    if (__builtin_expect(!(s <= 0x1p1000), 0)) {
        w0 = aa0 / s;
        w1 = aa1 / s;
        w2 = aa2 / s;
    }
}

// WENO-Z non-linear weights (+ AdaWENO indicator theta), weno_mod.f90:67-95,168-188
template <bool FAST>
__device__ __forceinline__ void weno5z_omega(const double* lf, double& w0, double& w1, double& w2, double* theta)
{
    double is0, is1, is2;
    smooth_ind<FAST>(lf, is0, is1, is2);
    const double tau5 = fabs(is2 - is0);
    double t;
    t = tau5 / (WENO_SS + is0);
    const double aa0 = C30 * (1.0 + t * t);
    t = tau5 / (WENO_SS + is1);
    const double aa1 = C31 * (1.0 + t * t);
    t = tau5 / (WENO_SS + is2);
    const double aa2 = C32 * (1.0 + t * t);
    const double s = aa0 + aa1 + aa2;
    normalize3<FAST>(aa0, aa1, aa2, s, w0, w1, w2);
    if (theta) {
        const double d = s - 1.0;
        *theta = 1.0 / (1.0 + d * d);
    }
}

// weno5z_stencil, weno_mod.f90:50-65
template <bool FAST>
__device__ __forceinline__ double weno5z(const double* lf)
{
    double q0, q1, q2, w0, w1, w2;
    candidates(lf, q0, q1, q2);
    weno5z_omega<FAST>(lf, w0, w1, w2, nullptr);
    return w0 * q0 + w1 * q1 + w2 * q2;
}

// weno5js_stencil, weno_mod.f90:98-165.  FAST: a_k = c_k/(eps+beta_k)^2 through
// reciprocals while (eps+beta_k)^2 <= 2^900 (then a_k in [2^-904, 2^40] and
// sum <= 2^42), '/' otherwise.
template <bool FAST>
__device__ __forceinline__ double weno5js(const double* lf)
{
    double q0, q1, q2, is0, is1, is2, s;
    candidates(lf, q0, q1, q2);
    smooth_ind<FAST>(lf, is0, is1, is2);
    s = WENO_SS + is0;
    const double d0 = s * s;
    s = WENO_SS + is1;
    const double d1 = s * s;
    s = WENO_SS + is2;
    const double d2 = s * s;
    double aa0, aa1, aa2;
    if constexpr (FAST) {
        aa0 = div_by(C30, d0, nr_recip(d0));
        aa1 = div_by(C31, d1, nr_recip(d1));
        aa2 = div_by(C32, d2, nr_recip(d2));
        if (__builtin_expect(!(d0 <= 0x1p900 && d1 <= 0x1p900 && d2 <= 0x1p900), 0)) {
            aa0 = C30 / d0;
            aa1 = C31 / d1;
            aa2 = C32 / d2;
        }
    } else {
        aa0 = C30 / d0;
        aa1 = C31 / d1;
        aa2 = C32 / d2;
    }
    const double sum = aa0 + aa1 + aa2;
    double w0, w1, w2;
    normalize3<FAST>(aa0, aa1, aa2, sum, w0, w1, w2);
    return w0 * q0 + w1 * q1 + w2 * q2;
}

// scheme_upwind5_general.f90:86-90
__device__ __forceinline__ double upwind5(const double* lf)
{
    return (2.0 * lf[0] - 13.0 * lf[1] + 47.0 * lf[2] + 27.0 * lf[3] - 3.0 * lf[4]) / 60.0;
}

template <int R>
__device__ __forceinline__ double recon(const double* lf)
{
    if constexpr (R == R_UP5) return upwind5(lf);
    else if constexpr (R == R_JS) return weno5js<false>(lf);
    else return weno5z<false>(lf);
}

// ---------------------------------------------------------------- cells

// Primitive variables of one cell, get_primary.f90:52-61.  pr = (u, v, p, c).
// WITH_C = false leaves out the sound speed (pr[3] = 0) for callers that
// never read it.
template <bool WITH_C = true>
__device__ __forceinline__ void prim_of(double gamma, const double* c, double* pr)
{
    const double rho = c[0];
    const double u = c[1] / rho;
    const double v = c[2] / rho;
    double p = (gamma - 1.0) * (c[3] - 0.5 * rho * (u * u + v * v));
    if (p < 0.0) p = P_FLOOR;
    pr[0] = u;
    pr[1] = v;
    pr[2] = p;
    pr[3] = WITH_C ? sqrt(gamma * p / rho) : 0.0;
}

// Dirichlet ghost from a primitive state, boundary_doublemach_general.f90:52-61
__device__ __forceinline__ void from_state(double gamma, double rho, double u, double v, double p, double* c,
                                           double* pr)
{
    c[0] = rho;
    c[1] = rho * u;
    c[2] = rho * v;
    c[3] = p / (gamma - 1.0) + 0.5 * rho * (u * u + v * v);
    pr[0] = u;
    pr[1] = v;
    pr[2] = p;
    pr[3] = sqrt(gamma * p / rho);
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
    const double* sv = a.bc.state[side];

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
        const double y = ((double)(gy + 1) - 0.5) * a.dy - sv[0];
        const double e = a.bc_rows[a.nj + gy];
        const double v = a.st->bc_t[0] * e + a.st->bc_t[1] * e;
        if (y >= 0.0) from_state(a.gamma, 1.6374, a.bc_rows[gy], v, 0.3327, c, pr);
        else from_state(a.gamma, 0.3626, a.bc_rows[gy], v, 0.3327, c, pr);
        return;
    }
    case DENSE_BC_DMR_TOP: {
        // boundary_doublemach_general.f90:67-80, shock at t^n
        const double x = ((double)(gx + 1) - 0.5) * a.dx;
        if (x < (1.0 / 6.0 + (1.0 + 20.0 * t) / a.sqrt3))
            from_state(a.gamma, sv[0], sv[1], sv[2], sv[3], c, pr);
        else
            from_state(a.gamma, 1.4, 0.0, 0.0, 1.0, c, pr);
        return;
    }
    default:  // DIRICHLET
        from_state(a.gamma, sv[0], sv[1], sv[2], sv[3], c, pr);
        return;
    }
}

// ---------------------------------------------------------------- fluxes

// Lax-Friedrichs split of one cell along DIR, sign +1/-1:
// F = 0.5*((u_n +/- alpha)*U + (0, p e_n, u_n p)),
// flux_splitting_LLF.f90:39-47 (x), :64-72 (y); GLF flux_splitting_GLF.f90:31-43.
template <int DIR>
__device__ __forceinline__ void lf_flux(double upa, double un, double p, const double* c, double* f)
{
    constexpr int NRM = DIR == 0 ? 1 : 2;
    f[0] = 0.5 * (upa * c[0]);
    f[NRM] = 0.5 * (upa * c[NRM] + p);
    f[3 - NRM] = 0.5 * (upa * c[3 - NRM]);
    f[3] = 0.5 * (upa * c[3] + un * p);
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
#pragma unroll
    for (int sgn = 0; sgn < 2; sgn++) {
        double* f = sgn == 0 ? fp : fm;
        double lp[3];
#pragma unroll
        for (int q = 0; q < 3; q++) {
            const double sq = sqrt(lam[q] * lam[q] + epsl * epsl);
            lp[q] = sgn == 0 ? (lam[q] + sq) / 2.0 : (lam[q] - sq) / 2.0;
        }
        const double wii = (3.0 - gamma) * (lp[1] + lp[2]) * (alpha * alpha) / (2.0 * (gamma - 1.0));
        f[0] = (lp[0] * 2.0 * (gamma - 1.0) + lp[1] + lp[2]) * rho / (2.0 * gamma);
        if (DIR == 0) {
            f[1] = (lp[0] * 2.0 * (gamma - 1.0) * u + lp[1] * (u - alpha) + lp[2] * (u + alpha)) * rho /
                   (2.0 * gamma);
            f[2] = (lp[0] * 2.0 * (gamma - 1.0) * v + lp[1] * v + lp[2] * v) * rho / (2.0 * gamma);
            f[3] = (lp[0] * (gamma - 1.0) * uu2 + lp[1] * 0.5 * ((u - alpha) * (u - alpha) + v * v) +
                    lp[2] * 0.5 * ((u + alpha) * (u + alpha) + v * v) + wii) *
                   rho / (2.0 * gamma);
        } else {
            f[1] = (lp[0] * 2.0 * (gamma - 1.0) * u + lp[1] * u + lp[2] * u) * rho / (2.0 * gamma);
            f[2] = (lp[0] * 2.0 * (gamma - 1.0) * v + lp[1] * (v - alpha) + lp[2] * (v + alpha)) * rho /
                   (2.0 * gamma);
            f[3] = (lp[0] * (gamma - 1.0) * uu2 + lp[1] * 0.5 * ((v - alpha) * (v - alpha) + u * u) +
                    lp[2] * 0.5 * ((v + alpha) * (v + alpha) + u * u) + wii) *
                   rho / (2.0 * gamma);
        }
    }
}

}  // namespace dense
