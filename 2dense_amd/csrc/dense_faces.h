// Face reconstructions of one tile.  A face thread reads the 6 cells of its
// stencil (f-3 .. f+2 along the sweep direction) from the LDS tile record and
// produces h(face) = h+(face) + h-(face) for the 4 conserved components,
// exactly as the reference's sweeps followed by sum_field produce h_xi/h_eta.
#pragma once
#include "dense_device.h"

namespace dense {

// The characteristic faces (switch 104 and 105) reconstruct their two flux
// signs one after the other at 3 waves per SIMD: 104 at 166 VGPRs with
// 20 B/lane of spill, 72 us per stage launch on double Mach 1920x480,
// against 90 us with the signs interleaved at 2 waves (243 VGPRs); AdaWENO
// 200 vs 215 us on Rayleigh-Taylor 1024x4096 and 63 vs 69 us on double Mach
// at t = 0.1, where serial signs also let each sign's indicator alone decide
// whether the face's Roe state is built.

// LDS record slots of one cell: 0..3 conserved (rho, rho u, rho v, E), then u, v, p, c.
enum { REC_U = 4, REC_V = 5, REC_P = 6, REC_C = 7, REC_N = 8 };

// Component-wise reconstruction with Lax-Friedrichs splitting.
//   LLF: alpha = |u_n|+c per cell (flux_splitting_LLF.f90:37);
//   GLF: alpha = global max (flux_splitting_GLF.f90:31-43).
// Stencils: + side lf(m) = F+(f-1+m), - side lf(m) = F-(f-m)
// (scheme_weno5z_general.f90:81, :106).
template <int DIR, bool LOCAL, int RECON, int NC, int ST>
__device__ __forceinline__ void face_cw_lf(const double* __restrict__ s, int base, double alpha, double* h)
{
    constexpr int NRM = DIR == 0 ? 1 : 2;
    constexpr int VN = DIR == 0 ? REC_U : REC_V;
    double up[6], um[6], pp[6], unp[6];
#pragma unroll
    for (int m = 0; m < 6; m++) {
        const int o = base + m * ST;
        const double un = s[VN * NC + o], p = s[REC_P * NC + o];
        const double a = LOCAL ? fabs(un) + s[REC_C * NC + o] : alpha;
        up[m] = un + a;
        um[m] = un - a;
        pp[m] = p;
        unp[m] = un * p;
    }
#pragma unroll
    for (int k = 0; k < 4; k++) {
        double lfp[5], lfm[5];
#pragma unroll
        for (int i = 0; i < 5; i++) {
            const int ip = i, im = 5 - i;
            const double qp = s[k * NC + base + ip * ST], qm = s[k * NC + base + im * ST];
            if (k == NRM) {
                lfp[i] = 0.5 * (up[ip] * qp + pp[ip]);
                lfm[i] = 0.5 * (um[im] * qm + pp[im]);
            } else if (k == 3) {
                lfp[i] = 0.5 * (up[ip] * qp + unp[ip]);
                lfm[i] = 0.5 * (um[im] * qm + unp[im]);
            } else {
                lfp[i] = 0.5 * (up[ip] * qp);
                lfm[i] = 0.5 * (um[im] * qm);
            }
        }
        h[k] = recon<RECON>(lfp) + recon<RECON>(lfm);
    }
}

// Component-wise reconstruction with Steger-Warming splitting.
template <int DIR, int RECON, int NC, int ST>
__device__ __forceinline__ void face_cw_sw(const double* __restrict__ s, int base, double gamma, double* h)
{
    double fp[6][4], fm[6][4];
#pragma unroll
    for (int m = 0; m < 6; m++) {
        const int o = base + m * ST;
        sw_flux<DIR>(gamma, s[o], s[REC_U * NC + o], s[REC_V * NC + o], s[REC_C * NC + o], fp[m], fm[m]);
    }
#pragma unroll
    for (int k = 0; k < 4; k++) {
        double lfp[5], lfm[5];
#pragma unroll
        for (int i = 0; i < 5; i++) {
            lfp[i] = fp[i][k];
            lfm[i] = fm[5 - i][k];
        }
        h[k] = recon<RECON>(lfp) + recon<RECON>(lfm);
    }
}

// ---- pieces of the characteristic faces (switch 104 and 105)
//   Roe average of cells f-1, f      scheme_weno5th_CH.f90:44-62
//   eigenvectors L, R                scheme_weno5th_SC.f90:277-328 (x), :607-658 (y)
//   projection, WENO-Z, back-proj.   scheme_weno5th_CH.f90:64-81
//   AdaWENO indicator rho+F_n, theta scheme_weno5th_SC.f90:1-64, flux_splitting_SC.f90:58-59
//   shared-weight branch (theta>delta) scheme_weno5th_SC.f90:110-124
// The eigenvectors are written in the frame (normal, tangential); the Roe
// average and the eigenvector entries share the reciprocals of rho_L, rho_R
// and 1 + sqrt(rho_L/rho_R), and the Roe sound speed and its reciprocal come
// from one reciprocal square root.

// Roe state and left eigenvectors of one face, halved (they project doubled
// split fluxes, lf_flux2).  Chars 1, 2 and 4 have entries for (rho, n, t, E)
// in Lr/Ln/Lt/Le[0..2]; char 3 is (-u_t, [0|1], [1|0], 0), halved in char_recon.
struct RoeFace {
    double Lr[3], Ln[3], Lt[3], Le[3];
    double ut0, un0, a0, h0, uu0;
};

template <int DIR, int NC, int ST>
__device__ __forceinline__ void roe_face(const double* __restrict__ s, int base, double gamma, RoeFace& R)
{
    const int oL = base + 2 * ST, oR = base + 3 * ST;
    const double rhol = s[oL], ul = s[REC_U * NC + oL], vl = s[REC_V * NC + oL], pl = s[REC_P * NC + oL];
    const double rhor = s[oR], ur = s[REC_U * NC + oR], vr = s[REC_V * NC + oR], pr = s[REC_P * NC + oR];
    // enthalpy H = p/(gamma-1)/rho + (u^2+v^2)/2 + p/rho = gamma/(gamma-1) p/rho + (u^2+v^2)/2
    const double gm1 = gamma - 1.0, gg1 = gamma * rcp(gm1);
    const double irl = rcp(rhol), irr = rcp(rhor);
    const double hl = gg1 * (pl * irl) + 0.5 * (ul * ul + vl * vl);
    const double hr = gg1 * (pr * irr) + 0.5 * (ur * ur + vr * vr);
    // Roe weights sqrt(rho_R)/(sqrt(rho_R)+sqrt(rho_L)) = 1/(1+t) and t/(1+t),
    // t = sqrt(rho_L/rho_R): one square root instead of two
    double t, rt;
    sqrt_rsq(rhol * irr, t, rt);
    const double wr = rcp(1.0 + t), wl = t * wr;
    const double u0 = wr * ur + wl * ul;
    const double v0 = wr * vr + wl * vl;
    R.h0 = wr * hr + wl * hl;
    double ia;  // a0 and 1/a0 from one reciprocal square root
    sqrt_rsq(gm1 * (R.h0 - 0.5 * (u0 * u0 + v0 * v0)), R.a0, ia);
    R.un0 = DIR == 0 ? u0 : v0;
    R.ut0 = DIR == 0 ? v0 : u0;
    R.uu0 = 0.5 * (u0 * u0 + v0 * v0);
    // 1/hh0 = (gamma-1)/a0^2, with hh0 = a0^2/(gamma-1) (scheme_weno5th_SC.f90:277-328)
    const double ihh = gm1 * (ia * ia);
    // entries of the eigenvectors times 0.5 (exact): 0.5 -> 0.25 in the rows of
    // chars 1 and 4, (1, u_n, u_t, -1) -> (0.5, 0.5 u_n, 0.5 u_t, -0.5) in char 2
    const double A = 0.25 * R.uu0 * ihh, B = 0.25 * R.un0 * ia, Cc = 0.25 * R.un0 * ihh, D = 0.25 * ia;
    const double E = 0.25 * R.ut0 * ihh, F = 0.25 * ihh;
    R.Lr[0] = A + B;  R.Ln[0] = -Cc - D;  R.Lt[0] = -E;  R.Le[0] = F;
    R.Lr[1] = 0.5 - (0.5 * R.uu0) * ihh;  R.Ln[1] = (0.5 * R.un0) * ihh;  R.Lt[1] = (0.5 * R.ut0) * ihh;
    R.Le[1] = -0.5 * ihh;
    R.Lr[2] = A - B;  R.Ln[2] = -Cc + D;  R.Lt[2] = -E;  R.Le[2] = F;
}

// The 5-point split-flux stencil of one sign, doubled (lf_flux2): + side
// 2F+(f-3+i), - side 2F-(f+2-i), i = 0..4 (scheme_weno5z_general.f90:81, :106
// with GLF fluxes).
template <int DIR, int NC, int ST>
__device__ __forceinline__ void flux_stencil(const double* __restrict__ s, int base, bool plus, double alpha,
                                             double F[5][4])
{
    constexpr int VN = DIR == 0 ? REC_U : REC_V;
    const double sa = plus ? alpha : -alpha;
#pragma unroll
    for (int i = 0; i < 5; i++) {
        const int o = base + (plus ? i : 5 - i) * ST;
        const double un = s[VN * NC + o];
        const double upa = un + sa;
        const double c4[4] = {s[o], s[NC + o], s[2 * NC + o], s[3 * NC + o]};
        lf_flux2<DIR>(upa, un, s[REC_P * NC + o], c4, F[i]);
    }
}

// Characteristic reconstruction of one sign: project the doubled stencil on
// the halved left eigenvectors (the reference's projections bit for bit),
// WENO-Z per characteristic field, back-project with R.
template <int DIR>
__device__ __forceinline__ void char_recon(const double F[5][4], const RoeFace& R, double* r4)
{
    constexpr int NRM = DIR == 0 ? 1 : 2, TAN = 3 - NRM;
    double acc_r = 0, acc_n = 0, acc_t = 0, acc_e = 0;
#pragma unroll
    for (int kc = 0; kc < 4; kc++) {
        double lf[5];
        if (kc == 2) {
            // l(:,3) = (-u_t, [0|1], [1|0], 0): ((-u_t)*2F_rho + 2F_t) * 0.5
#pragma unroll
            for (int i = 0; i < 5; i++) lf[i] = ((-R.ut0) * F[i][0] + F[i][TAN]) * 0.5;
        } else {
            const int q = kc == 3 ? 2 : kc;
            double l4[4];
            l4[0] = R.Lr[q];
            l4[NRM] = R.Ln[q];
            l4[TAN] = R.Lt[q];
            l4[3] = R.Le[q];
#pragma unroll
            for (int i = 0; i < 5; i++)
                lf[i] = l4[0] * F[i][0] + l4[1] * F[i][1] + l4[2] * F[i][2] + l4[3] * F[i][3];
        }
        const double cv = weno5z(lf);
        // back-projection h(c) = sum_k r(k,c) cv_k, accumulated in k order
        if (kc == 0) {
            acc_r = cv;
            acc_n = (R.un0 - R.a0) * cv;
            acc_t = R.ut0 * cv;
            acc_e = (R.h0 - R.a0 * R.un0) * cv;
        } else if (kc == 1) {
            acc_r = acc_r + cv;
            acc_n = acc_n + R.un0 * cv;
            acc_t = acc_t + R.ut0 * cv;
            acc_e = acc_e + R.uu0 * cv;
        } else if (kc == 2) {
            acc_t = acc_t + cv;
            acc_e = acc_e + R.ut0 * cv;
        } else {
            acc_r = acc_r + cv;
            acc_n = acc_n + (R.un0 + R.a0) * cv;
            acc_t = acc_t + R.ut0 * cv;
            acc_e = acc_e + (R.h0 + R.a0 * R.un0) * cv;
        }
    }
    r4[0] = acc_r;
    r4[NRM] = acc_n;
    r4[TAN] = acc_t;
    r4[3] = acc_e;
}

// AdaWENO's shared-weight branch of one sign: every component of the doubled
// stencil reconstructed with the indicator's WENO-Z weights, halved
// (scheme_weno5th_SC.f90:111-124).
__device__ __forceinline__ void cw_shared(const double F[5][4], double w0, double w1, double w2, double* r4)
{
#pragma unroll
    for (int k = 0; k < 4; k++) {
        double lf[5], q0, q1, q2;
#pragma unroll
        for (int i = 0; i < 5; i++) lf[i] = F[i][k];
        candidates(lf, q0, q1, q2);
        r4[k] = w0 * q0 + w1 * q1 + w2 * q2;
    }
}

// AdaWENO's indicator of one sign: WENO-Z weights of rho + F_n on the sign's
// stencil, and the branch: characteristic unless theta > delta (strict), with
// theta = 1/(1 + (S - 1)^2), S the sum of the unnormalised WENO-Z weights
// (weno_mod.f90:94, scheme_weno5th_SC.f90:24-28, :110).  For 0 < delta < 1,
// theta > delta is (S - 1)^2 < (1 - delta)/delta; with the weights scaled by
// P^2 = (d0 d1 d2)^2 as in weno5z (S' = S P^2) that is
// (S' - P^2)^2 < (1 - delta)/delta P^4: no quotient, and one reciprocal for
// the shared weights w_k = a'_k / S'.  theta <= 1 always, so delta >= 1 takes
// the characteristic branch everywhere and delta <= 0 the shared-weight one.
// The weights returned are halved (cw_shared applies them to a doubled stencil).
template <int DIR, int NC, int ST>
__device__ __forceinline__ bool adaweno_sign(const double* __restrict__ s, int base, bool plus, double alpha,
                                            double delta, double& w0, double& w1, double& w2)
{
    constexpr int NRM = DIR == 0 ? 1 : 2;
    constexpr int VN = DIR == 0 ? REC_U : REC_V;
    double lfi[5];
    const double sa = plus ? alpha : -alpha;
#pragma unroll
    for (int i = 0; i < 5; i++) {
        const int o = base + (plus ? i : 5 - i) * ST;
        const double un = s[VN * NC + o];
        lfi[i] = s[o] + 0.5 * ((un + sa) * s[NRM * NC + o] + s[REC_P * NC + o]);
    }
    double is0, is1, is2;
    smooth_ind(lfi, is0, is1, is2);
    const double tau5 = fabs(is2 - is0), t2 = tau5 * tau5;
    const double d0 = WENO_SS + is0, d1 = WENO_SS + is1, d2 = WENO_SS + is2;
    const double e0 = d1 * d2, e1 = d0 * d2, e2 = d0 * d1;
    const double aa0 = C30 * (d0 * d0 + t2) * (e0 * e0);
    const double aa1 = C31 * (d1 * d1 + t2) * (e1 * e1);
    const double aa2 = C32 * (d2 * d2 + t2) * (e2 * e2);
    const double sw = aa0 + aa1 + aa2;
    const double p = e0 * d0, p2 = p * p;
    bool shared;
    if (delta >= 1.0) shared = false;
    else if (delta <= 0.0) shared = true;
    else {
        const double k = (1.0 - delta) * rcp(delta);
        const double m = sw - p2;
        shared = m * m < k * (p2 * p2);
    }
    if (shared) {
        const double ys = 0.5 * rcp(sw);  // halved: the weights multiply candidates of 2F
        w0 = aa0 * ys;
        w1 = aa1 * ys;
        w2 = aa2 * ys;
    }
    return !shared;
}

// Characteristic WENO-Z (switch 104) at one face: h(face) = h+ + h- into
// h[0..3] (h[4..7] scratch), the signs one after the other.
template <int DIR, int NC, int ST>
__device__ __forceinline__ void face_char(const double* __restrict__ s, int base, double alpha, double gamma,
                                          double* h)
{
    RoeFace R;
    roe_face<DIR, NC, ST>(s, base, gamma, R);
#pragma unroll 1
    for (int sg = 0; sg < 2; sg++) {
        const bool plus = sg == 0;
        double F[5][4], r4[4];
        flux_stencil<DIR, NC, ST>(s, base, plus, alpha, F);
        char_recon<DIR>(F, R, r4);
        if (plus) {
#pragma unroll
            for (int k = 0; k < 4; k++) h[k] = r4[k];
        } else {
#pragma unroll
            for (int k = 0; k < 4; k++) h[4 + k] = r4[k];
        }
    }
#pragma unroll
    for (int k = 0; k < 4; k++) h[k] = h[k] + h[4 + k];
}

// AdaWENO (switch 105) at one face, the signs one after the other: each
// sign's indicator decides its branch (scheme_weno5th_SC.f90:110), and the
// Roe state is built at the first sign that takes the characteristic branch
// (once per face at most; most faces of developed flows never need it).
template <int DIR, int NC, int ST>
__device__ __forceinline__ void face_ada(const double* __restrict__ s, int base, double alpha, double gamma,
                                         double delta, double* h)
{
    RoeFace R;
    bool have_roe = false;
#pragma unroll 1
    for (int sg = 0; sg < 2; sg++) {
        const bool plus = sg == 0;
        double w0 = 0, w1 = 0, w2 = 0;
        const bool ch = adaweno_sign<DIR, NC, ST>(s, base, plus, alpha, delta, w0, w1, w2);
        double F[5][4], r4[4];
        flux_stencil<DIR, NC, ST>(s, base, plus, alpha, F);
        if (ch) {
            if (!have_roe) {
                roe_face<DIR, NC, ST>(s, base, gamma, R);
                have_roe = true;
            }
            char_recon<DIR>(F, R, r4);
        } else {
            cw_shared(F, w0, w1, w2, r4);
        }
        if (plus) {
#pragma unroll
            for (int k = 0; k < 4; k++) h[k] = r4[k];
        } else {
#pragma unroll
            for (int k = 0; k < 4; k++) h[4 + k] = r4[k];
        }
    }
#pragma unroll
    for (int k = 0; k < 4; k++) h[k] = h[k] + h[4 + k];
}

}  // namespace dense
