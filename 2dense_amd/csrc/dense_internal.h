// Internal structures shared by the kernels and the C-ABI layer.
#pragma once
#include <hip/hip_runtime.h>
#include <cstddef>
#include <cstdint>
#include "../../include/dense_capi.h"

namespace dense {

constexpr int BF = 4;  // ghost layers of the reference layout

enum Scheme : int {
    S_LLF_UP5 = 0, S_LLF_JS, S_LLF_Z,
    S_GLF_UP5, S_GLF_JS, S_GLF_Z,
    S_SW_UP5, S_SW_JS, S_SW_Z,
    S_CH, S_SC,
    S_COUNT
};

// Map the reference's (splitting, recon) switch pair to a kernel scheme, or -1.
inline int scheme_of(int splitting, int recon)
{
    int r = recon == 251 ? 0 : recon == 253 ? 1 : recon == 255 ? 2 : -1;
    switch (splitting) {
    case 101: return r < 0 ? -1 : S_LLF_UP5 + r;
    case 102: return r < 0 ? -1 : S_GLF_UP5 + r;
    case 103: return r < 0 ? -1 : S_SW_UP5 + r;
    case 104: return S_CH;  // always WENO-Z (scheme_weno5th_CH.f90:73)
    case 105: return S_SC;
    default: return -1;
    }
}

// A counter that every workgroup of a launch updates atomically.  Each sits on
// its own 128-byte line, so that the atomics on one counter do not serialise
// with those on another or with the clock words every workgroup reads at start
// (device-scope atomics on one line serialise: measured 39 us for 512
// workgroups x 8 queue fetches on one line, 19 us on separate lines).
struct alignas(128) LaunchCounter {
    unsigned int v;
};

// Device-resident clock and reduction slots of one solver (or one batch instance).
struct alignas(128) DevState {
    unsigned long long alpha[4][2];  // ring of max(|u|+c), max(|v|+c) bit patterns
    alignas(128) double t;           // t^n (the clock words: one line, read by every workgroup)
    double dt;                       // step size of the current step
    double t_end;                    // run while t < t_end ...
    double t_prev;                   // t^n of the last completed step (its stages' boundary time)
    int step;                        // ... and step < step_end
    int step_end;
    int seq;                         // kernel sequence number: reads alpha[seq&3], writes alpha[(seq+1)&3]
    int active;                      // 1 while the current step runs
    double elasp;                    // the save clock: sum of dt since the last Tecplot save ...
    double elasp_end;                // ... run while !(elasp >= elasp_end) (general_main.f90:37,40)
    LaunchCounter blocks_done;       // last-block detection
    LaunchCounter qnext[8];          // per-XCD tile queues of the persistent stage kernel
    int error;                       // 1: non-finite wave speed
    int hist_n;                      // dt history entries written
    int hist_cap;
    double bc_t[2];                  // clock terms of the shear-layer inflow at t^n (set by the host)
};

// Boundary rules of the four sides.  The ghost records (conserved state, then
// u, v, p, c) are derived from the primitive states once per configuration by
// bc_fill, so the stage kernels read them instead of dividing: per side the
// record of state[side] (Dirichlet sides, the double-Mach wall's Dirichlet
// part, the post-shock state of the moving-shock top), the pre-shock state
// (1.4, 0, 0, 1) of the moving-shock top, and for the shear-layer inflow
// p/(gamma-1) and the sound speeds of its two densities.
struct BcDesc {
    int kind[4];
    double state[4][4];  // rho, u, v, p
    double rec[4][8];
    double pre[8];
    double ssl_e, ssl_c[2];
};

// Shear-layer inflow (boundary_ssl_general.f90:55-85): the two densities
// (y >= 0, y < 0) and the common pressure.
constexpr double SSL_RHO_HI = 1.6374, SSL_RHO_LO = 0.3626, SSL_P = 0.3327;

// Ghost record of a primitive state (boundary_doublemach_general.f90:52-61).
inline void ghost_record(double gamma, double rho, double u, double v, double p, double* r)
{
    r[0] = rho;
    r[1] = rho * u;
    r[2] = rho * v;
    r[3] = p / (gamma - 1.0) + 0.5 * rho * (u * u + v * v);
    r[4] = u;
    r[5] = v;
    r[6] = p;
    r[7] = __builtin_sqrt(gamma * p / rho);
}

inline void bc_fill(BcDesc& b, const int kind[4], const double state[4][4], double gamma)
{
    for (int s = 0; s < 4; s++) {
        b.kind[s] = kind[s];
        for (int v = 0; v < 4; v++) b.state[s][v] = state[s][v];
        ghost_record(gamma, state[s][0], state[s][1], state[s][2], state[s][3], b.rec[s]);
    }
    ghost_record(gamma, 1.4, 0.0, 0.0, 1.0, b.pre);
    b.ssl_e = SSL_P / (gamma - 1.0);
    b.ssl_c[0] = __builtin_sqrt(gamma * SSL_P / SSL_RHO_HI);
    b.ssl_c[1] = __builtin_sqrt(gamma * SSL_P / SSL_RHO_LO);
}

// Arguments of one RK-stage launch (passed by value).
struct StageArgs {
    const double* un;   // U^n      (4 planes, pitch ld, nj rows)
    const double* us;   // stage input U^(s-1)
    double* out;        // stage output
    DevState* st;
    double* dt_hist;
    int ni, nj, ld;
    long long plane;    // ld * nj
    double dx, dy, gamma, cfl, delta, dt_const, sqrt3;
    int if_const_dt, if_bodyforce;
    const double* dt_sched;  // if set: dt of step k is dt_sched[k] (frozen schedule)
    int stage;          // 1..3
    int last_stage;
    int tiles_x, ntiles;
    BcDesc bc;
    const double* bc_rows;  // shear-layer inflow: u(y_j) for j < nj, then exp(-y_j^2/10)
};

// ---- batched mode: many independent blocks in one launch per RK stage
struct BatchInst {          // static description of one instance (device array)
    double* U[3];           // U^n, stage 1, stage 2 (same layout as a single block)
    int ni, nj, ld, tiles_x;
    long long plane;
    int tile_base, ntiles;  // its tiles in the batch's tile list
    double dx, dy, gamma, cfl, delta, dt_const;
    int if_const_dt, if_bodyforce;
    BcDesc bc;
};
struct alignas(64) BatchInstState {  // per-instance clock and reduction slots
    unsigned long long alpha[4][2];
    double t, dt;
    int step, error;
};
struct alignas(128) BatchState {     // batch-wide launch bookkeeping
    int seq, active, step, step_end;
    LaunchCounter blocks_done;
    LaunchCounter qnext[8];
};
struct BatchArgs {
    const BatchInst* inst;
    BatchInstState* ist;
    BatchState* bst;
    const int* tile_inst;   // global tile -> instance
    double* dt_hist;        // [step][instance], hist_cap steps
    int hist_cap;
    int ninst, ntiles;
    int stage, last_stage, nstages;
    double sqrt3;
};

hipError_t launch_batch_stage(int scheme, int tile, int grid_pct, const BatchArgs& b, hipStream_t s);
hipError_t launch_batch_max(const BatchArgs& b, hipStream_t s);

// host-side launchers (dense_kernels.hip)
// AdaWENO's indicator theta at every x and y face of the state a.us (boundary
// time t, global wave speeds ai, aj) into the reference's theta arrays.
hipError_t launch_theta(const StageArgs& a, double ai, double aj, double t, double* theta, hipStream_t s);
hipError_t launch_stage(int scheme, int tile, const StageArgs& a, hipStream_t s);
hipError_t launch_max(const StageArgs& a, hipStream_t s);
// Set the device clock (t, step), the stop condition and the save clock
// (elasp, elasp_end; infinite: no save stop), clear the dt history count.
hipError_t launch_set_clock(DevState* st, double t, double t_end, int step, int step_end, hipStream_t s,
                            double bc_t0 = 0.0, double bc_t1 = 0.0, double elasp = 0.0,
                            double elasp_end = __builtin_inf());
// Padded reference-layout cons(ni+8,nj+8,4) and prim(ni+8,nj+8,5): interior
// cells from a.us, ghost layers from the boundary rules applied to ghost_src
// at time t_ghost (corners zero).
hipError_t launch_fill_padded(const StageArgs& a, const double* ghost_src, double t_ghost, double* cons_pad,
                              double* prim_pad, hipStream_t s);
int tile_count(int tile, int ni, int nj, int* tiles_x);

// dense_cfg.tile -> index of the tile shape list (dense_tile.h: 0 16x15,
// 1 32x7, 2 8x7, 3 32x15).  0 (automatic) is the 16x15 tile: timed on MI355X
// for splittings 101, 102, 104 and 105 on 256^2 to 3840x960 grids, it is the
// fastest shape or within 5 % of it, except for GLF + upwind5 on the smallest
// grids (32x15 is 4-12 % faster there); the 32x15 tile is up to 31 % slower on
// large grids (two 256-thread workgroups per CU hide each other's load and
// barrier phases; one 512-thread workgroup cannot).  dense_autotune times the
// shapes for a given grid.  1..3 force those shapes, 4 forces 16x15 and 5 the
// 16x15 tile with 512 threads, one face per thread (shape 4 of the list:
// faster by a few % on some grids of 40,000-230,000 cells, slower on large ones).
inline int resolve_tile(int tile, int /*splitting*/, long long /*cells*/)
{
    return tile == 4 ? 0 : tile == 5 ? 4 : tile;
}

}  // namespace dense
