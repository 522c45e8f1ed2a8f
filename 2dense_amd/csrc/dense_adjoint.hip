// Hand-written FP64 discrete adjoint of the explicit update for
// Lax-Friedrichs splitting (101) + WENO-Z (255) + SSP-RK3 on a periodic grid.
//
// The forward map of one step (time_advance_general.f90:35-217) is
//   U1 = U0 + dt L(U0);  U2 = 3/4 U0 + 1/4 U1 + 1/4 dt L(U1);
//   U3 = 1/3 U0 + 2/3 U2 + 2/3 dt L(U2),
//   L(U)_i = -(h_x(i+1/2) - h_x(i-1/2))/dx - (h_y(j+1/2) - h_y(j-1/2))/dy,
// with the step sizes frozen (recorded by the forward pass), so the reverse
// pass is the exact transpose of the discrete map:
//   g2 = 2/3 g3 + 2/3 dt J(U2)^T g3,  g1 = 1/4 g2 + 1/4 dt J(U1)^T g2,
//   g0 = g1 + dt J(U0)^T g1 + 1/3 g3 + 3/4 g2.
// J^T g is applied by one tiled kernel per stage (adj_tile_kernel): it
// back-propagates lambda_h(face) = c dt (g(f) - g(f-1))/dx through the WENO-Z
// stencils of both flux signs of every component, gathers the stencil
// adjoints per cell in a fixed order (deterministic) and applies the transpose
// of the Lax-Friedrichs flux Jacobian, all in LDS.
// Stage states are recomputed segment by segment from checkpoints of U^n taken
// every `k` steps, with the forward stage kernel (bit-identical to the forward pass).
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <new>
#include <vector>

#include "dense_device.h"

namespace dense {

int validate_cfg(const dense_cfg* c, char* why, size_t n);

// Reverse-mode derivative of WENO-Z: lam_lf[m] += lam_h * dh/dlf[m].  The
// gradient need not reproduce the forward rounding, so quotients by a shared
// divisor go through one reciprocal.
__device__ __forceinline__ void weno5z_vjp(const double* lf, double lam_h, double* lam_lf)
{
    const double q0 = A300 * lf[0] + A301 * lf[1] + A302 * lf[2];
    const double q1 = A310 * lf[1] + A311 * lf[2] + A312 * lf[3];
    const double q2 = A320 * lf[2] + A321 * lf[3] + A322 * lf[4];
    const double X0 = (lf[0] - 2.0 * lf[1]) + lf[2], Y0 = (lf[0] - 4.0 * lf[1]) + 3.0 * lf[2];
    const double X1 = (lf[1] - 2.0 * lf[2]) + lf[3], Y1 = lf[1] - lf[3];
    const double X2 = (lf[2] - 2.0 * lf[3]) + lf[4], Y2 = (3.0 * lf[2] - 4.0 * lf[3]) + lf[4];
    constexpr double K13_12 = 13.0 / 12.0;
    const double b0 = K13_12 * (X0 * X0) + 0.25 * (Y0 * Y0);
    const double b1 = K13_12 * (X1 * X1) + 0.25 * (Y1 * Y1);
    const double b2 = K13_12 * (X2 * X2) + 0.25 * (Y2 * Y2);
    const double d = b2 - b0, tau = fabs(d);
    const double ie0 = 1.0 / (WENO_SS + b0), ie1 = 1.0 / (WENO_SS + b1), ie2 = 1.0 / (WENO_SS + b2);
    const double t0 = tau * ie0, t1 = tau * ie1, t2 = tau * ie2;
    const double a0 = C30 * (1.0 + t0 * t0), a1 = C31 * (1.0 + t1 * t1), a2 = C32 * (1.0 + t2 * t2);
    const double is = 1.0 / (a0 + a1 + a2);
    const double w0 = a0 * is, w1 = a1 * is, w2 = a2 * is;
    // reverse
    const double lq0 = lam_h * w0, lq1 = lam_h * w1, lq2 = lam_h * w2;
    const double lw0 = lam_h * q0, lw1 = lam_h * q1, lw2 = lam_h * q2;
    const double dot = lw0 * w0 + lw1 * w1 + lw2 * w2;
    const double la0 = (lw0 - dot) * is, la1 = (lw1 - dot) * is, la2 = (lw2 - dot) * is;
    const double lt0 = la0 * (2.0 * C30) * t0, lt1 = la1 * (2.0 * C31) * t1, lt2 = la2 * (2.0 * C32) * t2;
    const double ltau = lt0 * ie0 + lt1 * ie1 + lt2 * ie2;
    double lb0 = -lt0 * t0 * ie0, lb1 = -lt1 * t1 * ie1, lb2 = -lt2 * t2 * ie2;
    const double sg = d > 0.0 ? 1.0 : (d < 0.0 ? -1.0 : 0.0);
    lb2 += ltau * sg;
    lb0 -= ltau * sg;
    const double lX0 = lb0 * (13.0 / 6.0) * X0, lY0 = lb0 * 0.5 * Y0;
    const double lX1 = lb1 * (13.0 / 6.0) * X1, lY1 = lb1 * 0.5 * Y1;
    const double lX2 = lb2 * (13.0 / 6.0) * X2, lY2 = lb2 * 0.5 * Y2;
    lam_lf[0] += lX0 + lY0 + A300 * lq0;
    lam_lf[1] += -2.0 * lX0 - 4.0 * lY0 + lX1 + lY1 + A301 * lq0 + A310 * lq1;
    lam_lf[2] += lX0 + 3.0 * lY0 - 2.0 * lX1 + lX2 + 3.0 * lY2 + A302 * lq0 + A311 * lq1 + A320 * lq2;
    lam_lf[3] += lX1 - lY1 - 2.0 * lX2 - 4.0 * lY2 + A312 * lq1 + A321 * lq2;
    lam_lf[4] += lX2 + lY2 + A322 * lq2;
}

// Primitive variables of one cell for the derivative kernels (reciprocal of rho
// shared; the forward rounding need not be reproduced).
struct CellPrim {
    double ir, u, v, p, c;
    bool floored;  // pressure floor active: p does not depend on U
};

__device__ __forceinline__ CellPrim cell_prim(double gamma, const double* U)
{
    CellPrim q;
    q.ir = 1.0 / U[0];
    q.u = U[1] * q.ir;
    q.v = U[2] * q.ir;
    double p = (gamma - 1.0) * (U[3] - 0.5 * U[0] * (q.u * q.u + q.v * q.v));
    q.floored = p < 0.0;
    q.p = q.floored ? P_FLOOR : p;
    q.c = sqrt(gamma * q.p * q.ir);
    return q;
}

// Transpose of the Lax-Friedrichs flux Jacobian of one cell, direction DIR,
// sign sg (+1/-1): lam_U += (dF/dU)^T lam_F, F = 0.5*((u_n + sg*(|u_n|+c))*U + (0, p e_n, u_n p)).
template <int DIR>
__device__ __forceinline__ void lf_flux_vjp(double gamma, const double* U, const CellPrim& q, double sg,
                                            const double* lF, double* lU)
{
    constexpr int NRM = DIR == 0 ? 1 : 2, TAN = 3 - NRM;
    const double un = DIR == 0 ? q.u : q.v;
    const double w = un + sg * (fabs(un) + q.c);
    // F0 = .5 w rho, Fn = .5 (w mn + p), Ft = .5 w mt, FE = .5 (w E + un p)
    const double lw = 0.5 * (lF[0] * U[0] + lF[NRM] * U[NRM] + lF[TAN] * U[TAN] + lF[3] * U[3]);
    double l[4];
    l[0] = 0.5 * w * lF[0];
    l[NRM] = 0.5 * w * lF[NRM];
    l[TAN] = 0.5 * w * lF[TAN];
    l[3] = 0.5 * w * lF[3];
    double lp = 0.5 * lF[NRM] + 0.5 * un * lF[3];
    double lun = 0.5 * q.p * lF[3];
    const double sgn_un = un > 0.0 ? 1.0 : (un < 0.0 ? -1.0 : 0.0);
    lun += lw * (1.0 + sg * sgn_un);
    const double lc = sg * lw;
    // c = sqrt(gamma p / rho): dc/dp = c/(2p), dc/drho = -c/(2 rho)
    const double hc = 0.5 * lc * q.c;
    lp += hc / q.p;
    l[0] -= hc * q.ir;
    // u_n = m_n / rho
    l[NRM] += lun * q.ir;
    l[0] -= un * lun * q.ir;
    // p = (gamma-1)(E - (mx^2 + my^2)/(2 rho))
    if (!q.floored) {
        const double g1 = gamma - 1.0;
        l[3] += g1 * lp;
        l[1] -= g1 * q.u * lp;
        l[2] -= g1 * q.v * lp;
        l[0] += g1 * 0.5 * (q.u * q.u + q.v * q.v) * lp;
    }
#pragma unroll
    for (int k = 0; k < 4; k++) lU[k] += l[k];
}

struct AdjArgs {
    const double* U;   // stage input state, 4 planes of pitch ld
    const double* g;   // adjoint of the stage output
    const double* acc_in;
    double* g_out;
    double* acc_out;
    int ni, nj, ld;
    long long plane;
    double gamma, cdt_dx, cdt_dy;  // c*dt/dx, c*dt/dy of the stage
    double bcoef, acoef;           // out = bcoef*g + c dt J^T g (+ acc_in if add_acc); acc_out = acc_in + acoef*g
    int use_acc_in, add_acc, write_acc;
};

__device__ __forceinline__ int wrapi(int i, int n)
{
    i %= n;
    return i < 0 ? i + n : i;
}

// One workgroup of ANT threads per tile of ATX x ATY cells applies c dt J(U)^T g
// to it, in LDS.  Strips of the tile with a 4-cell periodic halo (x strip: rows of the
// tile, columns x0-4 .. x0+ATX+3; y strip: columns of the tile, rows
// y0-4 .. y0+ATY+3) hold U, u, v, p, c and g of each cell.  Four rounds, one
// per (direction, flux sign) in the order x+, x-, y+, y-:
//  * faces: one thread per face whose stencil reaches a tile cell.
//    lambda_h(f) = c dt (g(f) - g(f-1)) / dx is back-propagated through the
//    WENO-Z stencil of that sign for each component (lf(m) = F+(f-3+m) or
//    F-(f+2-m)); the 5 stencil adjoints per component go to LDS;
//  * cells: each tile cell sums the 5 contributions to its split flux from
//    the faces whose stencils hold it (a fixed order: deterministic) and
//    applies the transpose of the Lax-Friedrichs split-flux Jacobian.
// The four parts are summed in round order and the RK combination applied.
constexpr int ATX = 16, ATY = 8, ANT = 512;  // 512 threads: 1.5 face items per thread and round (measured 9 % faster than 256)
constexpr int AXW = ATX + 8, AYH = ATY + 8;          // strip extents
constexpr int AXF = (ATX + 4) * ATY;                 // x faces per round (row-major, ATX+4 per row)
constexpr int AYF = ATX * (ATY + 4);                 // y faces per round (ATX per face row)
constexpr int AMAXF = AXF > AYF ? AXF : AYF;
constexpr int ASX = AXW * ATY, ASY = ATX * AYH;      // strip cells
static_assert(ATX * ATY <= ANT, "one cell per thread");

struct AdjLds {
    // strip records, slot-major: U0..U3, u, v, p, c, g0..g3
    double sx[12][ASX];
    double sy[12][ASY];
    double fb[20][AMAXF];  // stencil adjoints of the round's faces: [m*4 + k][face]
};

template <int DIR>
__device__ __forceinline__ void adj_face_round(const AdjArgs& a, AdjLds& L, int sign, int tid)
{
    constexpr int NF = DIR == 0 ? AXF : AYF;
    constexpr int NRM = DIR == 0 ? 1 : 2;
    auto S = [&](int slot, int c) -> double {
        if constexpr (DIR == 0) return L.sx[slot][c];
        else return L.sy[slot][c];
    };
    const double scale = DIR == 0 ? a.cdt_dx : a.cdt_dy;
    const double sg = sign ? -1.0 : 1.0;
    // work items (face, component): the 4 components of a face are independent
    for (int item = tid; item < 4 * NF; item += ANT) {
        const int k = item / NF, face = item - k * NF;
        // strip coordinate along DIR of cell f (face f lies between cells f-1 and f)
        // and the other coordinate
        int pos, lane;
        if (DIR == 0) {
            const int row = face / (ATX + 4), fi = face - row * (ATX + 4);
            pos = fi + (sign ? 2 : 3);  // f = x0 + fi - (sign ? 2 : 1); strip column of f = f - x0 + 4
            lane = row;
        } else {
            const int fr = face / ATX, col = face - fr * ATX;
            pos = fr + (sign ? 2 : 3);
            lane = col;
        }
        auto at = [&](int p) { return DIR == 0 ? lane * AXW + p : p * ATX + lane; };
        const double lam_h = scale * (S(8 + k, at(pos)) - S(8 + k, at(pos - 1)));
        double lf[5], ll[5] = {0, 0, 0, 0, 0};
#pragma unroll
        for (int m = 0; m < 5; m++) {
            const int c = at(pos + (sign ? 2 - m : m - 3));
            const double un = S(4 + DIR, c);
            const double w = un + sg * (fabs(un) + S(7, c));
            const double ext = k == NRM ? S(6, c) : (k == 3 ? un * S(6, c) : 0.0);
            lf[m] = 0.5 * (w * S(k, c) + ext);
        }
        weno5z_vjp(lf, lam_h, ll);
#pragma unroll
        for (int m = 0; m < 5; m++) L.fb[m * 4 + k][face] = ll[m];
    }
}

// The split-flux part of tile cell (lx, ly) for one round.
template <int DIR>
__device__ __forceinline__ void adj_cell_round(const AdjArgs& a, const AdjLds& L, int sign, int lx, int ly,
                                               const double* U, const CellPrim& cp, double* part)
{
    double lF[4];
#pragma unroll
    for (int k = 0; k < 4; k++) {
        double sum = 0.0;
#pragma unroll
        for (int m = 0; m < 5; m++) {
            // + sign: the cell is lf(m) of face i+3-m; - sign: of face i-2+m
            int fidx;
            if (DIR == 0) fidx = ly * (ATX + 4) + (sign ? lx + m : lx + 4 - m);
            else fidx = (sign ? ly + m : ly + 4 - m) * ATX + lx;
            sum += L.fb[m * 4 + k][fidx];
        }
        lF[k] = sum;
    }
    lf_flux_vjp<DIR>(a.gamma, U, cp, sign ? -1.0 : 1.0, lF, part);
}

__global__ void __launch_bounds__(ANT) adj_tile_kernel(AdjArgs a)
{
    __shared__ AdjLds L;
    const int tid = threadIdx.x;
    const int ni = a.ni, nj = a.nj;
    const int x0 = blockIdx.x * ATX, y0 = blockIdx.y * ATY;
    // strips: U, primitives (the same cell_prim for every use of a cell), g
    for (int e = tid; e < ASX + ASY; e += ANT) {
        int gx, gy, q;
        double(*S)[ASX] = nullptr;
        double(*T)[ASY] = nullptr;
        if (e < ASX) {
            const int row = e / AXW, col = e - row * AXW;
            gx = wrapi(x0 + col - 4, ni);
            gy = wrapi(y0 + row, nj);
            S = L.sx;
            q = e;
        } else {
            const int r = e - ASX, row = r / ATX, col = r - row * ATX;
            gx = wrapi(x0 + col, ni);
            gy = wrapi(y0 + row - 4, nj);
            T = L.sy;
            q = r;
        }
        const long long o = (long long)gy * a.ld + gx;
        double U[4], G[4];
#pragma unroll
        for (int k = 0; k < 4; k++) {
            U[k] = a.U[k * a.plane + o];
            G[k] = a.g[k * a.plane + o];
        }
        const CellPrim cp = cell_prim(a.gamma, U);
        const double rec[12] = {U[0], U[1], U[2], U[3], cp.u, cp.v, cp.p, cp.c, G[0], G[1], G[2], G[3]};
#pragma unroll
        for (int s = 0; s < 12; s++) {
            if (S) S[s][q] = rec[s];
            else T[s][q] = rec[s];
        }
    }
    __syncthreads();

    // this thread's tile cell (threads beyond the tile or the grid have none)
    const int lx = tid % ATX, ly = tid / ATX;
    const bool live = tid < ATX * ATY && x0 + lx < ni && y0 + ly < nj;
    double U[4] = {1.0, 0.0, 0.0, 1.0};
    CellPrim cp{};
    if (live) {
        const int c = ly * AXW + lx + 4;
#pragma unroll
        for (int k = 0; k < 4; k++) U[k] = L.sx[k][c];
        cp = cell_prim(a.gamma, U);
    }
    double l[4] = {0, 0, 0, 0};
#pragma unroll
    for (int round = 0; round < 4; round++) {
        const int sign = round & 1;
        if (round < 2) adj_face_round<0>(a, L, sign, tid);
        else adj_face_round<1>(a, L, sign, tid);
        __syncthreads();
        if (live) {
            double part[4] = {0, 0, 0, 0};
            if (round < 2) adj_cell_round<0>(a, L, sign, lx, ly, U, cp, part);
            else adj_cell_round<1>(a, L, sign, lx, ly, U, cp, part);
#pragma unroll
            for (int k = 0; k < 4; k++) l[k] = round == 0 ? part[k] : l[k] + part[k];
        }
        __syncthreads();  // the next round rewrites the face buffer
    }
    if (!live) return;
    const long long o = (long long)(y0 + ly) * a.ld + (x0 + lx);
#pragma unroll
    for (int k = 0; k < 4; k++) {
        const double g = L.sx[8 + k][ly * AXW + lx + 4];
        const double acc = a.use_acc_in ? a.acc_in[k * a.plane + o] : 0.0;
        a.g_out[k * a.plane + o] = a.add_acc ? a.bcoef * g + l[k] + acc : a.bcoef * g + l[k];
        if (a.write_acc) a.acc_out[k * a.plane + o] = acc + a.acoef * g;
    }
}

// fresh wave-speed ring and queues before a forward or recomputed segment
__global__ void adj_reset_kernel(DevState* st, int hist_cap)
{
    for (int s = 0; s < 4; s++) st->alpha[s][0] = st->alpha[s][1] = 0ull;
    for (int x = 0; x < 8; x++) st->qnext[x].v = 0u;
    st->seq = 0;
    st->blocks_done.v = 0u;
    st->error = 0;
    st->active = 0;
    st->hist_cap = hist_cap;
}

}  // namespace dense

using namespace dense;

struct dense_adjoint {
    dense_cfg cfg{};
    int scheme = -1, nsteps = 0, k = 0, nseg = 0;
    hipStream_t stream = nullptr;   // forward pass and the reverse sweep
    hipStream_t rstream = nullptr;  // recomputation of the next segment, overlapped with the reverse sweep
    int ld = 0, tile = 0, tiles_x = 0, ntiles = 0;
    long long plane = 0;
    DevState* st = nullptr;
    double* dt_hist = nullptr;
    double* d_dts = nullptr;
    std::vector<double> h_dts;
    double* U = nullptr;     // forward work buffers: 3 states
    double* ckpt = nullptr;  // nseg states
    double* seg[2] = {nullptr, nullptr};  // stage states (3k+1 each) of two segments: one reversed, one recomputed
    double* gb = nullptr;    // g, g_next, acc (3 states)
    hipEvent_t ev[4] = {nullptr, nullptr, nullptr, nullptr};
    hipEvent_t ready[2] = {nullptr, nullptr}, freed[2] = {nullptr, nullptr}, join = nullptr;
    double fwd_ms = 0.0, bwd_ms = 0.0;
    bool have_forward = false;
    bool last_seg_fresh = false;  // seg[(nseg-1)&1] holds the last segment's stage states from the forward pass
    char err[512] = {0};
};

static thread_local char g_aerr[512];

static int afail(dense_adjoint* a, int code, const char* fmt, ...)
{
    char* buf = a ? a->err : g_aerr;
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, 512, fmt, ap);
    va_end(ap);
    return code;
}

#define ACHK(a, call)                                                                              \
    do {                                                                                           \
        hipError_t e_ = (call);                                                                    \
        if (e_ != hipSuccess) return afail(a, DENSE_ERR_HIP, "%s: %s", #call, hipGetErrorString(e_)); \
    } while (0)

static double* state_at(double* base, const dense_adjoint* a, long long i) { return base + i * 4 * a->plane; }

static StageArgs adj_stage_args(const dense_adjoint* a)
{
    StageArgs s{};
    const dense_cfg& c = a->cfg;
    s.st = a->st;
    s.dt_hist = a->dt_hist;
    s.ni = c.ni;
    s.nj = c.nj;
    s.ld = a->ld;
    s.plane = a->plane;
    s.dx = c.dx;
    s.dy = c.dy;
    s.gamma = c.gamma;
    s.cfl = c.cfl;
    s.delta = c.delta;
    s.dt_const = c.dt;
    s.sqrt3 = std::sqrt(3.0);
    s.if_const_dt = c.if_const_dt;
    s.tiles_x = a->tiles_x;
    s.ntiles = a->ntiles;
    bc_fill(s.bc, c.bc, c.bc_state, c.gamma);
    return s;
}

// one forward step U0 -> (U1, U2) -> U3 with explicit buffers
static int adj_enqueue_step(dense_adjoint* a, hipStream_t strm, double* U0, double* U1, double* U2, double* U3,
                            const double* sched)
{
    StageArgs s = adj_stage_args(a);
    s.dt_sched = sched;
    double* outs[3] = {U1, U2, U3};
    const double* ins[3] = {U0, U1, U2};
    for (int k = 1; k <= 3; k++) {
        s.stage = k;
        s.last_stage = (k == 3);
        s.un = U0;
        s.us = ins[k - 1];
        s.out = outs[k - 1];
        ACHK(a, launch_stage(a->scheme, a->tile, s, strm));
    }
    return DENSE_OK;
}

// fresh reduction ring + wave speeds of U0 (the stage kernels read them)
static int adj_prime(dense_adjoint* a, hipStream_t strm, double* U0, double t, int step, int step_end)
{
    hipLaunchKernelGGL(adj_reset_kernel, dim3(1), dim3(1), 0, strm, a->st, 4096);
    ACHK(a, hipGetLastError());
    StageArgs s = adj_stage_args(a);
    s.us = U0;
    ACHK(a, launch_max(s, strm));
    ACHK(a, launch_set_clock(a->st, t, INFINITY, step, step_end, strm));
    return DENSE_OK;
}

static void seg_range(const dense_adjoint* a, int sgi, int& n0, int& len)
{
    n0 = sgi * a->k;
    len = (n0 + a->k <= a->nsteps) ? a->k : a->nsteps - n0;
}

// stage states of segment sgi into seg[sgi&1] from its checkpoint, on the recompute stream
static int adj_recompute(dense_adjoint* a, int sgi)
{
    const size_t st_bytes = sizeof(double) * 4 * a->plane;
    double* S = a->seg[sgi & 1];
    int n0, len;
    seg_range(a, sgi, n0, len);
    ACHK(a, hipMemcpyAsync(S, state_at(a->ckpt, a, sgi), st_bytes, hipMemcpyDeviceToDevice, a->rstream));
    int rc = adj_prime(a, a->rstream, S, 0.0, n0, n0 + len);
    if (rc) return rc;
    for (int j = 0; j < len; j++)
        if ((rc = adj_enqueue_step(a, a->rstream, state_at(S, a, 3 * j), state_at(S, a, 3 * j + 1),
                                   state_at(S, a, 3 * j + 2), state_at(S, a, 3 * j + 3), a->d_dts)))
            return rc;
    return DENSE_OK;
}

extern "C" int dense_adj_create(const dense_cfg* cfg, int nsteps, int checkpoint_every, dense_adj_handle* out)
{
    if (!cfg || !out || nsteps < 1 || checkpoint_every < 1) return afail(nullptr, DENSE_ERR_ARG, "bad arguments");
    *out = nullptr;
    dense_cfg c = *cfg;
    char why[256];
    int rc = validate_cfg(&c, why, sizeof why);
    if (rc) return afail(nullptr, rc, "%s", why);
    if (c.splitting != 101 || c.recon != 255 || c.stages != 3)
        return afail(nullptr, DENSE_ERR_UNSUPPORTED, "the adjoint covers LLF (101) + WENO-Z (255) + RK3");
    for (int q = 0; q < 4; q++)
        if (c.bc[q] != DENSE_BC_PERIODIC) return afail(nullptr, DENSE_ERR_UNSUPPORTED, "the adjoint needs periodic boundaries");
    if (c.if_bodyforce) return afail(nullptr, DENSE_ERR_UNSUPPORTED, "no body force in the adjoint");
    if (nsteps > 4096) return afail(nullptr, DENSE_ERR_ARG, "at most 4096 steps");
    dense_adjoint* a = new (std::nothrow) dense_adjoint;
    if (!a) return afail(nullptr, DENSE_ERR_NOMEM, "host allocation failed");
    auto bail = [&](int code) {
        dense_adj_destroy(a);
        return code;
    };
    a->cfg = c;
    a->scheme = scheme_of(c.splitting, c.recon);
    a->nsteps = nsteps;
    a->k = checkpoint_every < nsteps ? checkpoint_every : nsteps;
    a->nseg = (nsteps + a->k - 1) / a->k;
    a->ld = (c.ni + 31) / 32 * 32;
    a->plane = (long long)a->ld * c.nj;
    a->tile = resolve_tile(c.tile, c.splitting, (long long)c.ni * c.nj);
    a->ntiles = tile_count(a->tile, c.ni, c.nj, &a->tiles_x);
    const size_t st_bytes = sizeof(double) * 4 * a->plane;
    if (hipStreamCreateWithFlags(&a->stream, hipStreamNonBlocking) != hipSuccess ||
        hipStreamCreateWithFlags(&a->rstream, hipStreamNonBlocking) != hipSuccess)
        return bail(afail(nullptr, DENSE_ERR_HIP, "stream creation failed"));
    for (auto& e : a->ev)
        if (hipEventCreate(&e) != hipSuccess) return bail(afail(nullptr, DENSE_ERR_HIP, "event creation failed"));
    for (int b = 0; b < 2; b++)
        if (hipEventCreateWithFlags(&a->ready[b], hipEventDisableTiming) != hipSuccess ||
            hipEventCreateWithFlags(&a->freed[b], hipEventDisableTiming) != hipSuccess)
            return bail(afail(nullptr, DENSE_ERR_HIP, "event creation failed"));
    if (hipEventCreateWithFlags(&a->join, hipEventDisableTiming) != hipSuccess)
        return bail(afail(nullptr, DENSE_ERR_HIP, "event creation failed"));
    const size_t seg_bytes = (size_t)(3 * a->k + 1) * st_bytes;
    if (hipMalloc(&a->st, sizeof(DevState)) != hipSuccess || hipMalloc(&a->dt_hist, sizeof(double) * 4096) != hipSuccess ||
        hipMalloc(&a->d_dts, sizeof(double) * nsteps) != hipSuccess || hipMalloc(&a->U, 3 * st_bytes) != hipSuccess ||
        hipMalloc(&a->ckpt, (size_t)a->nseg * st_bytes) != hipSuccess ||
        hipMalloc(&a->seg[0], seg_bytes) != hipSuccess || hipMalloc(&a->seg[1], seg_bytes) != hipSuccess ||
        hipMalloc(&a->gb, 3 * st_bytes) != hipSuccess)
        return bail(afail(nullptr, DENSE_ERR_NOMEM, "adjoint allocation failed"));
    if (hipMemset(a->U, 0, 3 * st_bytes) != hipSuccess || hipMemset(a->seg[0], 0, seg_bytes) != hipSuccess ||
        hipMemset(a->seg[1], 0, seg_bytes) != hipSuccess || hipMemset(a->gb, 0, 3 * st_bytes) != hipSuccess)
        return bail(afail(nullptr, DENSE_ERR_HIP, "memset failed"));
    a->h_dts.assign(nsteps, 0.0);
    *out = a;
    return DENSE_OK;
}

extern "C" int dense_adj_destroy(dense_adj_handle a)
{
    if (!a) return DENSE_OK;
    if (a->stream) (void)hipStreamSynchronize(a->stream);
    if (a->rstream) (void)hipStreamSynchronize(a->rstream);
    double* bufs[] = {a->dt_hist, a->d_dts, a->U, a->ckpt, a->seg[0], a->seg[1], a->gb};
    for (double* p : bufs)
        if (p) (void)hipFree(p);
    if (a->st) (void)hipFree(a->st);
    for (auto& e : a->ev)
        if (e) (void)hipEventDestroy(e);
    for (int b = 0; b < 2; b++) {
        if (a->ready[b]) (void)hipEventDestroy(a->ready[b]);
        if (a->freed[b]) (void)hipEventDestroy(a->freed[b]);
    }
    if (a->join) (void)hipEventDestroy(a->join);
    if (a->stream) (void)hipStreamDestroy(a->stream);
    if (a->rstream) (void)hipStreamDestroy(a->rstream);
    delete a;
    return DENSE_OK;
}

// copy between a packed (4, nj, ni) device array and a pitched state
static int copy_packed(dense_adjoint* a, double* dst, size_t dpitch, const double* src, size_t spitch)
{
    const dense_cfg& c = a->cfg;
    ACHK(a, hipMemcpy2DAsync(dst, dpitch * sizeof(double), src, spitch * sizeof(double), sizeof(double) * c.ni,
                             (size_t)4 * c.nj, hipMemcpyDeviceToDevice, a->stream));
    return DENSE_OK;
}

extern "C" int dense_adj_forward(dense_adj_handle a, const double* u0_dev, double* un_dev, const double* dt_schedule,
                                 double* dts_out)
{
    if (!a || !u0_dev || !un_dev) return afail(a, DENSE_ERR_ARG, "null argument");
    const dense_cfg& c = a->cfg;
    const size_t st_bytes = sizeof(double) * 4 * a->plane;
    a->have_forward = false;
    a->last_seg_fresh = false;
    int rc = copy_packed(a, a->U, a->ld, u0_dev, c.ni);
    if (rc) return rc;
    if (dt_schedule)
        ACHK(a, hipMemcpyAsync(a->d_dts, dt_schedule, sizeof(double) * a->nsteps, hipMemcpyHostToDevice, a->stream));
    if ((rc = adj_prime(a, a->stream, a->U, 0.0, 0, a->nsteps))) return rc;
    // the last segment's stage states are kept (in the buffer its reversal reads), the
    // others are recomputed from checkpoints in the backward pass
    const int last = a->nseg - 1;
    int ln0, llen;
    seg_range(a, last, ln0, llen);
    double* S = a->seg[last & 1];
    const double* sched = dt_schedule ? a->d_dts : nullptr;
    ACHK(a, hipEventRecord(a->ev[0], a->stream));
    for (int n = 0; n < a->nsteps; n++) {
        if (n % a->k == 0)
            ACHK(a, hipMemcpyAsync(state_at(a->ckpt, a, n / a->k), a->U, st_bytes,
                                   hipMemcpyDeviceToDevice, a->stream));
        if (n < ln0) {
            // stage 3 writes back in place (it reads U^n only at the cell it writes)
            rc = adj_enqueue_step(a, a->stream, a->U, state_at(a->U, a, 1), state_at(a->U, a, 2), a->U, sched);
        } else {
            const int j = n - ln0;
            if (j == 0) ACHK(a, hipMemcpyAsync(S, a->U, st_bytes, hipMemcpyDeviceToDevice, a->stream));
            rc = adj_enqueue_step(a, a->stream, state_at(S, a, 3 * j), state_at(S, a, 3 * j + 1),
                                  state_at(S, a, 3 * j + 2), state_at(S, a, 3 * j + 3), sched);
        }
        if (rc) return rc;
    }
    ACHK(a, hipEventRecord(a->ev[1], a->stream));
    if (!dt_schedule)
        ACHK(a, hipMemcpyAsync(a->d_dts, a->dt_hist, sizeof(double) * a->nsteps, hipMemcpyDeviceToDevice, a->stream));
    if ((rc = copy_packed(a, un_dev, c.ni, state_at(S, a, 3 * llen), a->ld))) return rc;
    ACHK(a, hipMemcpyAsync(a->h_dts.data(), a->d_dts, sizeof(double) * a->nsteps, hipMemcpyDeviceToHost, a->stream));
    DevState sh{};
    ACHK(a, hipMemcpyAsync(&sh, a->st, sizeof sh, hipMemcpyDeviceToHost, a->stream));
    ACHK(a, hipStreamSynchronize(a->stream));
    float ms = 0.f;
    ACHK(a, hipEventElapsedTime(&ms, a->ev[0], a->ev[1]));
    a->fwd_ms = ms;
    if (dts_out) std::memcpy(dts_out, a->h_dts.data(), sizeof(double) * a->nsteps);
    if (sh.error) return afail(a, DENSE_ERR_NONFINITE, "non-finite wave speed in the forward pass");
    a->have_forward = true;
    a->last_seg_fresh = true;
    return DENSE_OK;
}

static int adj_vjp(dense_adjoint* a, const double* Us, const double* g, double* gout, const double* acc_in,
                   double* acc_out, bool add_acc, double ccoef, double bcoef, double acoef, double dt)
{
    const dense_cfg& c = a->cfg;
    AdjArgs x{};
    x.U = Us;
    x.g = g;
    x.acc_in = acc_in;
    x.g_out = gout;
    x.acc_out = acc_out;
    x.ni = c.ni;
    x.nj = c.nj;
    x.ld = a->ld;
    x.plane = a->plane;
    x.gamma = c.gamma;
    x.cdt_dx = ccoef * dt / c.dx;
    x.cdt_dy = ccoef * dt / c.dy;
    x.bcoef = bcoef;
    x.acoef = acoef;
    x.use_acc_in = acc_in != nullptr;
    x.add_acc = add_acc;
    x.write_acc = acc_out != nullptr;
    const dim3 grid((unsigned)((c.ni + ATX - 1) / ATX), (unsigned)((c.nj + ATY - 1) / ATY));
    hipLaunchKernelGGL(adj_tile_kernel, grid, dim3(ANT), 0, a->stream, x);
    ACHK(a, hipGetLastError());
    return DENSE_OK;
}

extern "C" int dense_adj_backward(dense_adj_handle a, const double* gn_dev, double* g0_dev)
{
    if (!a || !gn_dev || !g0_dev) return afail(a, DENSE_ERR_ARG, "null argument");
    if (!a->have_forward) return afail(a, DENSE_ERR_ARG, "dense_adj_backward without a successful dense_adj_forward");
    const dense_cfg& c = a->cfg;
    double* g = state_at(a->gb, a, 0);
    double* gn = state_at(a->gb, a, 1);
    double* acc = state_at(a->gb, a, 2);
    int rc = copy_packed(a, g, a->ld, gn_dev, c.ni);
    if (rc) return rc;
    ACHK(a, hipEventRecord(a->ev[2], a->stream));
    ACHK(a, hipStreamWaitEvent(a->rstream, a->ev[2], 0));
    // Segment sgi is reversed on `stream` from seg[sgi&1] while segment sgi-1 is
    // recomputed on `rstream` into the other buffer; events order the hand-offs.
    // The last segment's states come from the forward pass (same stream) unless
    // a backward pass has run since.
    const int last = a->nseg - 1;
    bool used[2] = {false, false};
    const bool fresh = a->last_seg_fresh;
    if (!fresh) {
        if ((rc = adj_recompute(a, last))) return rc;
        ACHK(a, hipEventRecord(a->ready[last & 1], a->rstream));
    }
    for (int sgi = last; sgi >= 0; sgi--) {
        if (sgi > 0) {
            const int b = (sgi - 1) & 1;
            if (used[b]) ACHK(a, hipStreamWaitEvent(a->rstream, a->freed[b], 0));
            if ((rc = adj_recompute(a, sgi - 1))) return rc;
            ACHK(a, hipEventRecord(a->ready[b], a->rstream));
        }
        if (sgi != last || !fresh) ACHK(a, hipStreamWaitEvent(a->stream, a->ready[sgi & 1], 0));
        double* S = a->seg[sgi & 1];
        int n0, len;
        seg_range(a, sgi, n0, len);
        for (int j = len - 1; j >= 0; j--) {
            const double dt = a->h_dts[n0 + j];
            double* U0 = state_at(S, a, 3 * j);
            double* U1 = state_at(S, a, 3 * j + 1);
            double* U2 = state_at(S, a, 3 * j + 2);
            // stage 3: U3 = 1/3 U0 + 2/3 U2 + 2/3 dt L(U2)
            if ((rc = adj_vjp(a, U2, g, gn, nullptr, acc, false, 2.0 / 3.0, 2.0 / 3.0, 1.0 / 3.0, dt))) return rc;
            // stage 2: U2 = 3/4 U0 + 1/4 U1 + 1/4 dt L(U1)
            if ((rc = adj_vjp(a, U1, gn, g, acc, acc, false, 1.0 / 4.0, 1.0 / 4.0, 3.0 / 4.0, dt))) return rc;
            // stage 1: U1 = U0 + dt L(U0); g0 = g1 + dt J^T g1 + acc
            if ((rc = adj_vjp(a, U0, g, gn, acc, nullptr, true, 1.0, 1.0, 0.0, dt))) return rc;
            std::swap(g, gn);
        }
        ACHK(a, hipEventRecord(a->freed[sgi & 1], a->stream));
        used[sgi & 1] = true;
    }
    ACHK(a, hipEventRecord(a->join, a->rstream));
    ACHK(a, hipStreamWaitEvent(a->stream, a->join, 0));
    ACHK(a, hipEventRecord(a->ev[3], a->stream));
    a->last_seg_fresh = false;  // its buffer may have been reused
    if ((rc = copy_packed(a, g0_dev, c.ni, g, a->ld))) return rc;
    ACHK(a, hipStreamSynchronize(a->stream));
    float ms = 0.f;
    ACHK(a, hipEventElapsedTime(&ms, a->ev[2], a->ev[3]));
    a->bwd_ms = ms;
    return DENSE_OK;
}

extern "C" int dense_adj_last_ms(dense_adj_handle a, double* fwd_ms, double* bwd_ms)
{
    if (!a) return afail(a, DENSE_ERR_ARG, "null handle");
    if (fwd_ms) *fwd_ms = a->fwd_ms;
    if (bwd_ms) *bwd_ms = a->bwd_ms;
    return DENSE_OK;
}

extern "C" const char* dense_adj_last_error(dense_adj_handle a) { return a ? a->err : g_aerr; }
