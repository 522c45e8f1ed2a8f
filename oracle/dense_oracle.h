/*
 * dense_oracle.h -- CPU restatement of 2DENSE's explicit update.
 *
 * TEST INFRASTRUCTURE ONLY.  This is the checker that the HIP path is
 * compared against; only tests/, __graft_entry__.smoke() and bench.py's
 * cpu_baseline leg may load it.  The product library never links it.
 *
 * Layout follows the reference exactly (Fortran column-major, 4 ghost
 * layers): cons(ni+8, nj+8, 4), prim(ni+8, nj+8, 6), i fastest.
 */
#ifndef DENSE_ORACLE_H
#define DENSE_ORACLE_H

#ifdef __cplusplus
extern "C" {
#endif

/* Per-side boundary kinds (side order: 0 left/x-, 1 right/x+, 2 bottom/y-, 3 top/y+). */
enum {
    ORA_BC_PERIODIC = 0,
    ORA_BC_EXTRAP = 1,     /* 2nd-order extrapolation, cons and prim separately */
    ORA_BC_REFLECT = 2,    /* mirror with normal momentum / velocity negated */
    ORA_BC_DIRICHLET = 3,  /* fixed primitive state */
    ORA_BC_DMR_WALL = 4,   /* bottom of double Mach: Dirichlet post-shock for x<1/6, else reflect */
    ORA_BC_DMR_TOP = 5,    /* top of double Mach: moving shock at x = 1/6 + (1+20 t)/sqrt(3) */
    ORA_BC_SSL_INLET = 6   /* left of the shear layer: forced inflow; bc_state[0][0] = yl/2 */
};

typedef struct ora_params {
    int ni, nj;
    int splitting;      /* 101 LLF, 102 GLF, 103 SW, 104 CH, 105 SC */
    int recon;          /* 251, 253, 255 */
    int stages;         /* 1..3 */
    int if_const_dt;
    int if_bodyforce;
    double dx, dy;
    double gamma, cfl, delta, dt_const;
    int bc[4];                  /* side kinds */
    double bc_state[4][4];      /* Dirichlet primitive state (rho,u,v,p) per side */
    int nthreads;               /* OpenMP threads for the row loops (>=1) */
} ora_params;

typedef struct ora_block ora_block;

ora_block* ora_create(const ora_params* p);
void ora_destroy(ora_block* b);

/* Canned problems: fills params (grid metrics, BCs, gamma/bodyforce for RT)
 * and returns 0, or -1 for an unknown benchmark index.  Mirrors
 * init_general.f90:13-37 for indices 1, 3, 5, 9. */
int ora_problem_params(int benchmark, int ni, int nj, ora_params* p);
/* Fill the interior initial state of a created block (init_*.f90). */
int ora_init_state(ora_block* b, int benchmark);

/* Padded-array access: cons (ni+8)*(nj+8)*4, prim (ni+8)*(nj+8)*6. */
void ora_set_cons(ora_block* b, const double* cons);
void ora_get_cons(const ora_block* b, double* cons);
void ora_get_prim(const ora_block* b, double* prim);
void ora_set_bodyforce(ora_block* b, const double* bf);

/* One time step (general_main.f90:26-37): stages x {get_primary, boundary,
 * flux splitting + reconstruction, RK stage}, then step += 1, t += dt. */
void ora_step(ora_block* b);
/* Test-only: plant mistake m in the restatement (0 = none); see g_mut in dense_oracle.c. */
void ora_set_mutation(int m);
/* Test hooks: the Lax-Friedrichs and Steger-Warming splits of one cell
 * (nrm 2: x, 3: y); c = conserved state, a = wave speed used by LF. */
void ora_lf_split(double un, double p, double a, const double c[4], int nrm, double fp[4], double fn[4]);
void ora_sw_split(double gamma, double rho, double u, double v, double alpha, int nrm, double fp[4], double fn[4]);
/* Run while (t < total_time && step < total_step) (general_main.f90:23-24). */
int ora_run(ora_block* b, double total_time, int total_step);

double ora_time(const ora_block* b);
void ora_set_time(ora_block* b, double t);
int ora_current_step(const ora_block* b);
double ora_dt(const ora_block* b);
void ora_set_dt(ora_block* b, double dt);
void ora_max_eig(const ora_block* b, double out[2]);
void ora_update_params(ora_block* b, double gamma, double cfl, double delta,
                       int if_const_dt, double dt_const, int if_bodyforce);

/* Single-function probes used by the unit tests. */
double ora_weno5z_stencil(const double lf[5]);
double ora_weno5js_stencil(const double lf[5]);
double ora_upwind5_stencil(const double lf[5]);
void ora_weno5z_omega_theta(const double lf[5], double out[4]);
/* Eigenvectors of the x-split Euler Jacobian at a Roe state (l(comp,char), r(char,comp)),
 * column-major 4x4 as in the reference. */
void ora_lr_vec_xi(double u0, double v0, double h0, double a0, double gamma,
                   double l[16], double r[16]);
void ora_lr_vec_eta(double u0, double v0, double h0, double a0, double gamma,
                    double l[16], double r[16]);

/* Op-counting: number of +,-,*,/,sqrt executed by the last ora_step, by
 * the counting build (oracle compiled with -DORA_COUNT_OPS); -1 otherwise. */
long long ora_op_count(void);

#ifdef __cplusplus
}
#endif
#endif
