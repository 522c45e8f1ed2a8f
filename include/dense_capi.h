/*
 * dense_capi.h -- C ABI of the MI355X explicit-update library (lib2dense_amd.so).
 *
 * This is the drop-in boundary for 2DENSE's hot path.  In the reference the
 * hot path is four parameterless calls per Runge-Kutta stage on module state
 * (source/utility/general_main.f90:26-31):
 *     get_primary -> boundary_general -> flux_splitting_general -> time_rk3_general(irk)
 * A Fortran driver binds the functions below with iso_c_binding (module
 * 2dense_amd/fortran/dense_capi_mod.f90) and replaces those four calls with
 * one dense_step() per time step.
 *
 * Conventions
 *  - All arithmetic is IEEE FP64.  Host arrays use the reference layout:
 *    cons(ni+8, nj+8, 4) and prim(ni+8, nj+8, 6), Fortran column-major,
 *    i fastest, 4 ghost layers (flow_mod_general.f90:203-205).
 *  - Every function returns DENSE_OK (0) or a negative DENSE_ERR_* code;
 *    the library never prints, stops or aborts.  dense_last_error() gives text.
 *  - A handle is used by one host thread at a time; several handles may coexist.
 *  - The library owns all device memory; host arrays are owned by the caller.
 */
#ifndef DENSE_CAPI_H
#define DENSE_CAPI_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

#define DENSE_OK 0
#define DENSE_ERR_ARG (-1)         /* bad argument or size */
#define DENSE_ERR_UNSUPPORTED (-2) /* switch code without an implementation (the reference `stop`s) */
#define DENSE_ERR_HIP (-3)         /* HIP runtime error */
#define DENSE_ERR_NONFINITE (-4)   /* non-finite wave speed: the state has blown up */
#define DENSE_ERR_NOMEM (-5)       /* device allocation failed */

/* Boundary sides and kinds.  benchmark_index 1, 2, 3, 4, 5, 6, 9 select the reference's
 * own boundary routines (boundary_general.f90:5-31); benchmark 0 takes the
 * per-side description in dense_cfg.bc / bc_state. */
enum { DENSE_SIDE_LEFT = 0, DENSE_SIDE_RIGHT = 1, DENSE_SIDE_BOTTOM = 2, DENSE_SIDE_TOP = 3 };
enum {
    DENSE_BC_PERIODIC = 0,  /* boundary_ivc_general.f90:13-35, boundary_rt_general.f90:129-151 */
    DENSE_BC_EXTRAP = 1,    /* boundary_ivc_general.f90:61-102: (4 q1 - q2)/3, cons and prim separately */
    DENSE_BC_REFLECT = 2,   /* boundary_shockvortex_general.f90:15-79 */
    DENSE_BC_DIRICHLET = 3, /* boundary_rt_general.f90:14-57, boundary_doublemach_general.f90:98-118 */
    DENSE_BC_DMR_WALL = 4,  /* boundary_doublemach_general.f90:15-63 (bottom only) */
    DENSE_BC_DMR_TOP = 5,   /* boundary_doublemach_general.f90:65-96 (top only) */
    DENSE_BC_SSL_INLET = 6  /* boundary_ssl_general.f90:55-85, shear-layer inflow (left only);
                               bc_state[left][0] = yl/2, cell centre y = (j - 1/2) dy - yl/2 */
};

/* Run configuration: the NAMELIST variables the hot path reads
 * (flow_mod_general.f90:51-180) plus the grid metrics set by init_*. */
typedef struct dense_cfg {
    int ni, nj;              /* SINGLE_NI, SINGLE_NJ (interior cells) */
    int nghost;              /* bfsize_i = bfsize_j; must be 4 */
    int neq;                 /* NEQ; must be 4 */
    double dx, dy;           /* block%dx, block%dy */
    double gamma;            /* GAMMA */
    double cfl;              /* CFL */
    double delta;            /* DELTA (AdaWENO threshold) */
    double dt;               /* DT, used when if_const_dt != 0 */
    int splitting;           /* SCHEME_SPLITTING_INDEX: 101 LLF, 102 GLF, 103 SW, 104 CH, 105 AdaWENO */
    int recon;               /* SCHEME_INV_INDEX: 251 upwind5, 253 WENO-JS, 255 WENO-Z */
    int stages;              /* STAGES: 1, 2 or 3 */
    int if_const_dt;         /* IF_CONST_DT */
    int if_bodyforce;        /* gravity source of benchmark 3 (bodyforce = (0,0,rho,rho v)) */
    int benchmark;           /* BENCHMARK_INDEX: 1, 2, 3, 4, 5, 6, 9, or 0 for bc[] below */
    int bc[4];               /* per-side kind (benchmark 0) */
    double bc_state[4][4];   /* per-side Dirichlet (rho,u,v,p); DMR sides: post-shock state */
    int tile;                /* tile shape: 0 automatic (16x15), 1 32x7, 2 8x7, 3 32x15,
                                4 16x15, 5 16x15 with 512 threads (x and y faces on separate
                                halves); dense_autotune picks one by timing.  Results do not
                                depend on it. */
} dense_cfg;

typedef struct dense_solver* dense_handle;

/* Fill cfg->dx, dy, bc, bc_state (and gamma, if_bodyforce for benchmark 3) for
 * one of the reference's canned problems: init_general.f90:13-37.
 * Replaces the metric part of init_{IVC,sedov,RT,RM,doublemach,ssl,shockvortex}_general.f90. */
int dense_problem_defaults(int benchmark, int ni, int nj, dense_cfg* cfg);

/* Allocate device state for one block.
 * Replaces allocate_block (flow_mod_general.f90:190-295). */
int dense_create(const dense_cfg* cfg, dense_handle* out);

/* Free everything (the reference never frees). */
int dense_destroy(dense_handle h);

/* Upload the conserved state cons(ni+8,nj+8,4) (interior cells are read).
 * Replaces the module-state hand-off after init_general / load_date_binary
 * (init_general.f90:13-37, read_WriteAndLoad.f90:30-70). */
int dense_set_state(dense_handle h, const double* cons);

/* Download cons(ni+8,nj+8,4) and, if prim != NULL, prim(ni+8,nj+8,6).
 * Interior values are the current state; ghost layers are those the boundary
 * routines give the current state at the current time.  prim(:,:,6) is the
 * entropy log(p)-log(rho) on interior cells.  Replaces the get_primary call
 * before output (general_main.f90:45,60) and the arrays read by
 * write_tec_at_time / write_date_binary (output_test_general.f90:16-58,
 * read_WriteAndLoad.f90:1-28). */
int dense_get_state(dense_handle h, double* cons, double* prim);

/* Download cons(ni+8,nj+8,4) as the reference's array holds it when it writes
 * a .dat2d dump (write_date_binary, read_WriteAndLoad.f90:1-28, called at
 * general_main.f90:43-44): interior = the current state; ghost layers = the
 * reference's last boundary pass, i.e. the boundary rules applied to the last
 * stage's input state at t^n (boundary_general before time_rk3_general(stages),
 * general_main.f90:26-31); corner cells zero.  Before the first step after
 * dense_set_state or dense_autotune, the ghost layers of the current state at
 * the current time (as dense_get_state).  Stages = 1 keeps no copy of that
 * input: DENSE_ERR_UNSUPPORTED after a step. */
int dense_get_dump(dense_handle h, double* cons);

/* One time step: `stages` x {get_primary, boundary_general,
 * flux_splitting_general, time_rk3_general(irk)}
 * (general_main.f90:26-31).  current_time is t^n, used by time-dependent
 * boundaries in every stage (boundary_doublemach_general.f90:71).  *dt
 * receives the step size (CFL or constant, time_advance_general.f90:20-24);
 * the caller advances its own clock as general_main.f90:35-37 does. */
int dense_step(dense_handle h, double current_time, double* dt);
/* (The shear-layer inflow evaluates cos of t^n: with that boundary, dense_run
 * steps from the host, one synchronised step at a time, instead of replaying
 * graphs; results are the same.) */

/* Up to max_steps steps while (t < total_time && step < total_step)
 * (general_main.f90:23-24), replayed from captured HIP graphs (16-step chunks
 * and one graph of the remainder) with the clock on the device.  On return *t
 * and *step hold the advanced clock; dt_hist, if not NULL, must hold max_steps
 * entries and receives the dt of each step taken (*steps_taken of them). */
int dense_run(dense_handle h, int max_steps, double total_time, int total_step,
              double* t, int* step, double* dt_hist, int* steps_taken);

/* dense_run that also stops at the reference's save clock: *elasp is
 * elasp_time, advanced by dt after every step (general_main.f90:37), and the
 * run ends after the first step at which *elasp >= dt_save (the save test of
 * general_main.f90:40; at least one step runs, also when *elasp >= dt_save on
 * entry), as well as at total_time, total_step or max_steps.
 * A driver that replays the reference's loop calls it with max_steps = the
 * steps to the next multiple of SAVE_STEP, then saves when either test holds.
 * On return *elasp holds the advanced save clock. */
int dense_run_saving(dense_handle h, int max_steps, double total_time, int total_step, double dt_save,
                     double* elasp, double* t, int* step, double* dt_hist, int* steps_taken);

/* Allocation-free variant of dense_run.  No counterpart in the reference: it
 * is for callers that cannot allocate device memory -- code under HIP graph
 * capture, tight serving loops, harnesses that forbid hipMalloc.  It allocates
 * nothing: the run's scratch memory (device clock, launch counters, the two
 * stage buffers) is a fixed device array of the library (64 MiB + 4 KB per
 * GPU, held until the process ends), and no handle is created.
 *  u      DEVICE pointer to the block's interior state: 4 planes (rho, rho u,
 *         rho v, E) of nj rows with row pitch ld doubles (ld >= ni), advanced
 *         in place by nsteps steps (no early stop), each with its CFL step
 *         (or cfg->dt when if_const_dt) and boundary time t0 + the steps so far;
 *  stream a hipStream_t, or NULL for the null stream;
 *  t_out, dt_out  host pointers (nullable).  If either is given the call waits
 *         for the stream and returns the advanced clock and the last step size
 *         (and DENSE_ERR_NONFINITE for a blown-up state); if both are NULL it
 *         only enqueues kernels on the stream, so a caller may capture it.
 * Limit: ld * nj <= 1,048,576 (DMR 1920x480 fits); larger blocks return
 * DENSE_ERR_ARG.  The shear-layer inflow (host-evaluated clock terms) returns
 * DENSE_ERR_UNSUPPORTED.  Every call on one GPU shares that GPU's scratch and
 * leaves its contents behind: calls that overlap on one GPU (two host threads,
 * or two streams in flight) race.  Results are bit for bit those of dense_run
 * on the same state. */
int dense_run_noalloc(const dense_cfg* cfg, double* u, int ld, int nsteps, double t0, void* stream,
                      double* t_out, double* dt_out);

/* GPU time of the last dense_run, from HIP events recorded on the handle's
 * stream around the graph launches, and the number of RK-stage kernel
 * launches that did work in it.  No reference counterpart (the reference
 * times its loop with OMP_get_wtime, general_main.f90:25-33). */
int dense_last_run_gpu_ms(dense_handle h, double* ms, int* stage_launches);

/* Global wave speeds max|u|+c, max|v|+c of the current state (get_primary.f90:69-73). */
int dense_max_eig(dense_handle h, double out[2]);

/* Apply re-read NAMELIST values (gamma, cfl, delta, dt, if_const_dt,
 * if_bodyforce, splitting, recon) -- general_main.f90:53 re-reads input.in
 * at every save. */
int dense_update_params(dense_handle h, const dense_cfg* cfg);

/* Time each tile shape (cfg.tile 4, 3, 1, 2, 5) on trial_steps steps of the
 * current state, restoring state and clock after every trial, and keep the
 * fastest as an explicit cfg.tile (dense_update_params keeps it).  The trials
 * overwrite the stage buffers: dense_get_theta and the dump ghost layers of
 * dense_get_dump need a step after it.  *tile and
 * *best_ms (nullable) receive the choice and its trial time.  No counterpart in
 * the reference (one CPU code path); all shapes give the same bits. */
int dense_autotune(dense_handle h, int trial_steps, int* tile, double* best_ms);

/* AdaWENO's indicator theta (splitting 105, stages 2 or 3, after at least one
 * step) at every face, as the reference's theta arrays hold it when it writes
 * 2d_theta_<step>.dat (output_test_general.f90:61-119): from the last stage's
 * flux evaluation (get_theta_xi_p/_n, get_theta_eta_p/_n,
 * scheme_weno5th_SC.f90:1-64 and the eta analogues).  theta receives
 * theta_xip(ni+1,nj), theta_xin(ni+1,nj), theta_etap(nj+1,ni),
 * theta_etan(nj+1,ni), Fortran order, one after the other. */
int dense_get_theta(dense_handle h, double* theta);

/* Roofline calibration of the current device (no counterpart in the reference):
 * dense_probe_fp64 times a loop of independent FP64 FMAs over every CU and
 * returns the achieved TFLOP/s (the vector FP64 peak the stage kernels are
 * bound by); dense_probe_hbm times a streaming read and a streaming copy of
 * `bytes` bytes (>= 1 MiB; use sizes well beyond the 256 MB last-level cache)
 * and returns GB/s (the copy counts its reads and writes).  Best of 5 runs. */
int dense_probe_fp64(double* tflops);
int dense_probe_hbm(size_t bytes, double* read_gbs, double* copy_gbs);

/* Wait for queued device work of this handle. */
int dense_sync(dense_handle h);

/* Text of the last error on this handle (h may be NULL for create failures). */
const char* dense_last_error(dense_handle h);

/* ---------------------------------------------------------------------------
 * Batched mode (no counterpart in the reference, which advances one block):
 * n independent blocks of any sizes, same splitting/reconstruction/stages,
 * each with its own metrics, boundaries, gamma, CFL and clock, advanced by one
 * persistent launch per RK stage (graph-replayed).  Every instance evolves bit
 * for bit as it would through dense_step alone.  Needs stages >= 2.
 * ------------------------------------------------------------------------- */
typedef struct dense_batch* dense_batch_handle;

int dense_batch_create(const dense_cfg* cfgs, int n, dense_batch_handle* out);
int dense_batch_destroy(dense_batch_handle b);
/* Upload instance i's state cons(ni+8,nj+8,4) and its clock t0. */
int dense_batch_set_state(dense_batch_handle b, int i, const double* cons, double t0);
/* Download instance i's state (interior; ghost layers zero). */
int dense_batch_get_state(dense_batch_handle b, int i, double* cons);
/* nsteps steps of every instance, each with its own CFL dt.  dt_hist (nullable,
 * nsteps*n doubles) receives dt[step*n + i] for every step; t_out (nullable)
 * each instance's clock afterwards. */
int dense_batch_run(dense_batch_handle b, int nsteps, double* dt_hist, double* t_out);
/* Tile shape (0..3), instance order (0 as given, 1 largest first) and the
 * fraction of the resident capacity launched (1..100 %). */
int dense_batch_set_schedule(dense_batch_handle b, int tile, int order, int grid_pct);
/* Try every schedule on trial_steps steps of the current states (restored after
 * each trial) and keep the fastest. */
int dense_batch_autotune(dense_batch_handle b, int trial_steps, int* tile, int* order, int* grid_pct,
                         double* best_ms);
int dense_batch_last_run_gpu_ms(dense_batch_handle b, double* ms);
const char* dense_batch_last_error(dense_batch_handle b);

/* Initial states generated on the device, one descriptor per instance, for the
 * ensemble's two instance kinds.  The reference builds one state on the host
 * (init_IVC_general.f90:13-48, init_shockvortex_general.f90:13-57); these are
 * the same flows with the parameters free:
 *  - DENSE_IC_VORTEX: isentropic vortex of strength eps at (xc, yc) in the
 *    free stream rho = p = 1, (u, v) = (uinf, vinf), on the periodic square
 *    [0, length)^2 (minimum-image distance); density T^(1/(gamma-1)),
 *    T = 1 - (gamma-1) eps^2 e^(1-r^2) / (8 gamma pi^2), p = rho^gamma.
 *  - DENSE_IC_SHOCK_VORTEX: normal shock at rest at x = xs, upstream
 *    rho = p = 1, u = mach sqrt(gamma); downstream the Rankine-Hugoniot state
 *    (shock_relation.py:7-13); upstream a compressible vortex (eps, rc, alpha)
 *    centred at (xc, yc) as in init_shockvortex_general.f90:34-51.
 *  - DENSE_IC_SHOCK_VORTEX_REF: the reference's own benchmark-9 state,
 *    init_shockvortex_general.f90:34-56 operation for operation (its
 *    single-precision literals widened, Mach 1.1, the printed post-shock
 *    constants 1.169082, 1.113299, 0, 1.245); the descriptor's parameters are
 *    not read.  Equal to the reference's state bit for bit wherever the device
 *    exp and pow round as the host's do.
 * Cell centres are ((i + 1/2) dx, (j + 1/2) dy) with the instance's metrics.
 * Every instance's clock is reset to 0.  Replaces dense_batch_set_state for
 * all instances at once; no host state crosses the bus. */
enum { DENSE_IC_VORTEX = 1, DENSE_IC_SHOCK_VORTEX = 2, DENSE_IC_SHOCK_VORTEX_REF = 3 };
typedef struct dense_ic {
    int kind;            /* DENSE_IC_VORTEX or DENSE_IC_SHOCK_VORTEX */
    double eps;          /* vortex strength */
    double xc, yc;       /* vortex centre */
    double uinf, vinf;   /* free-stream velocity (vortex) */
    double length;       /* side of the periodic square (vortex) */
    double mach;         /* shock Mach number >= 1 (shock/vortex) */
    double xs;           /* shock position (shock/vortex) */
    double rc, alpha;    /* vortex core radius and decay rate (shock/vortex) */
} dense_ic;

int dense_batch_init_ic(dense_batch_handle b, const dense_ic* ics);

/* ---------------------------------------------------------------------------
 * Discrete adjoint (no counterpart in the reference, which has no gradients):
 * the exact transpose of nsteps steps of Lax-Friedrichs splitting (101) +
 * WENO-Z (255) + SSP-RK3 (time_advance_general.f90:35-217) on a periodic
 * block, with the step sizes frozen.  The forward pass keeps checkpoints of
 * U^n every checkpoint_every steps; the backward pass recomputes each
 * segment's stage states from its checkpoint with the forward kernels and
 * applies hand-written vector-Jacobian products.
 * States cross this interface as DEVICE pointers to packed float64 arrays
 * (4, nj, ni) (component, row, column; C order), so a caller that holds its
 * data in device memory (a torch tensor) never round-trips through the host.
 * ------------------------------------------------------------------------- */
typedef struct dense_adjoint* dense_adj_handle;

int dense_adj_create(const dense_cfg* cfg, int nsteps, int checkpoint_every, dense_adj_handle* out);
int dense_adj_destroy(dense_adj_handle a);
/* u0_dev -> un_dev over nsteps steps.  dt_schedule (host, nullable): frozen
 * step sizes; NULL takes the CFL step of every step (time_advance_general.f90:20-24).
 * dts_out (host, nullable) receives the nsteps step sizes used; the backward
 * pass uses the same ones. */
int dense_adj_forward(dense_adj_handle a, const double* u0_dev, double* un_dev, const double* dt_schedule,
                      double* dts_out);
/* g0_dev = (d un / d u0)^T gn_dev at the last forward pass. */
int dense_adj_backward(dense_adj_handle a, const double* gn_dev, double* g0_dev);
/* GPU time (HIP events) of the last forward and backward passes. */
int dense_adj_last_ms(dense_adj_handle a, double* fwd_ms, double* bwd_ms);
const char* dense_adj_last_error(dense_adj_handle a);

#ifdef __cplusplus
}
#endif
#endif
