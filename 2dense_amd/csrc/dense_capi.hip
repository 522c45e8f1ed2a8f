// C ABI of the explicit-update library: one solver handle per block, one HIP
// stream per handle, device-resident state (three interior buffers for the
// three RK stages, no ghost layers in HBM) and a device-resident clock.
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <atomic>
#include <climits>
#include <cstring>
#include <new>
#include <vector>

#include "dense_internal.h"

using namespace dense;

struct dense_solver {
    dense_cfg cfg{};
    int scheme = -1;
    hipStream_t stream = nullptr;
    int ld = 0;              // row pitch in doubles
    long long plane = 0;     // ld * nj
    double* U[3] = {nullptr, nullptr, nullptr};  // U^n, stage 1, stage 2
    DevState* st = nullptr;  // device
    DevState* st_host = nullptr;  // pinned readback
    double* dt_hist = nullptr;    // device
    int hist_cap = 0;
    double* pad_dev = nullptr;    // padded cons + prim for get_state
    double* theta_dev = nullptr;  // AdaWENO indicator at the faces, for get_theta
    bool stepped = false;         // a step ran since the state or the parameters were set
    bool stage_in_valid = false;  // U[stages-1] holds the last step's last-stage input
    double* pad_host = nullptr;   // pinned staging
    double* stage_host = nullptr; // pinned interior staging for set_state
    hipGraphExec_t graph = nullptr;      // GRAPH_CHUNK steps
    int graph_steps = 0;
    hipGraphExec_t graph_rem[16] = {};   // [r]: r < GRAPH_CHUNK steps, the remainder of a run
    int graph_rem_steps[16] = {};
    hipEvent_t ev0 = nullptr, ev1 = nullptr;  // bracket the last dense_run on the stream
    double last_run_ms = 0.0;
    int last_run_launches = 0;
    int tile = 0;            // resolved tile shape (resolve_tile of cfg.tile)
    int tiles_x = 0, ntiles = 0;
    double* bc_rows = nullptr;    // device: shear-layer inflow tables (2 nj), else null
    double* bct_host = nullptr;   // pinned: clock terms of the shear-layer inflow
    char err[512] = {0};
};

// Clock terms of the shear-layer inflow at t, with the host's libm as the
// reference evaluates them (boundary_ssl_general.f90:70-71):
// 0.05 cos(2 pi t/(30/2.68)) and 0.05 cos(2 pi 2 t/(30/2.68) + pi/2).
static void ssl_clock_terms(double t, double* out)
{
    const double pie = std::atan(1.0) * 4.0;
    out[0] = 0.05 * std::cos(2.0 * pie * t / (30.0 / 2.68));
    out[1] = 0.05 * std::cos(2.0 * pie * 2.0 * t / (30.0 / 2.68) + pie / 2.0);
}

static thread_local char g_err[512];

static int fail(dense_solver* h, int code, const char* fmt, ...)
{
    char* buf = h ? h->err : g_err;
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, 512, fmt, ap);
    va_end(ap);
    return code;
}

#define HIPCHK(h, call)                                                                        \
    do {                                                                                       \
        hipError_t e_ = (call);                                                                \
        if (e_ != hipSuccess) return fail(h, DENSE_ERR_HIP, "%s: %s", #call, hipGetErrorString(e_)); \
    } while (0)

namespace dense {
int validate_cfg(const dense_cfg* c, char* why, size_t n)
{
    if (c->ni < 4 || c->nj < 4) return snprintf(why, n, "ni, nj must be >= 4 (got %d, %d)", c->ni, c->nj), DENSE_ERR_ARG;
    if ((long long)c->ni * c->nj > (1ll << 31)) return snprintf(why, n, "grid too large"), DENSE_ERR_ARG;
    if (c->nghost != 4) return snprintf(why, n, "nghost must be 4 (reference bfsize)"), DENSE_ERR_ARG;
    if (c->neq != 4) return snprintf(why, n, "neq must be 4"), DENSE_ERR_ARG;
    if (!(c->dx > 0) || !(c->dy > 0)) return snprintf(why, n, "dx, dy must be positive"), DENSE_ERR_ARG;
    if (c->stages < 1 || c->stages > 3) return snprintf(why, n, "stages must be 1..3"), DENSE_ERR_ARG;
    if (scheme_of(c->splitting, c->recon) < 0)
        return snprintf(why, n, "unsupported switch pair splitting=%d recon=%d", c->splitting, c->recon),
               DENSE_ERR_UNSUPPORTED;
    if (c->tile < 0 || c->tile > 5) return snprintf(why, n, "tile index out of range"), DENSE_ERR_ARG;
    for (int s = 0; s < 4; s++)
        if (c->bc[s] < 0 || c->bc[s] > 6) return snprintf(why, n, "bad boundary kind on side %d", s), DENSE_ERR_ARG;
    for (int s = 1; s < 4; s++)
        if (c->bc[s] == DENSE_BC_SSL_INLET) return snprintf(why, n, "shear-layer inflow only on the left"), DENSE_ERR_ARG;
    if ((c->bc[0] == DENSE_BC_PERIODIC) != (c->bc[1] == DENSE_BC_PERIODIC) ||
        (c->bc[2] == DENSE_BC_PERIODIC) != (c->bc[3] == DENSE_BC_PERIODIC))
        return snprintf(why, n, "periodic boundaries must be paired"), DENSE_ERR_ARG;
    if (c->bc[0] == DENSE_BC_DMR_WALL || c->bc[0] == DENSE_BC_DMR_TOP || c->bc[1] == DENSE_BC_DMR_WALL ||
        c->bc[1] == DENSE_BC_DMR_TOP || c->bc[2] == DENSE_BC_DMR_TOP || c->bc[3] == DENSE_BC_DMR_WALL)
        return snprintf(why, n, "double-Mach wall/top kinds only on bottom/top"), DENSE_ERR_ARG;
    const int need = c->ni < c->nj ? c->ni : c->nj;
    if (need < 4) return snprintf(why, n, "grid must have at least 4 cells per direction"), DENSE_ERR_ARG;
    return DENSE_OK;
}
}  // namespace dense

extern "C" int dense_problem_defaults(int benchmark, int ni, int nj, dense_cfg* c)
{
    if (!c || ni < 2 || nj < 2) return fail(nullptr, DENSE_ERR_ARG, "bad arguments");
    c->ni = ni;
    c->nj = nj;
    c->nghost = 4;
    c->neq = 4;
    c->benchmark = benchmark;
    if (c->stages == 0) c->stages = 3;
    const double post[4] = {8.0, 8.25 * std::sqrt(3.0) * 0.5, -8.25 * 0.5, 116.5};
    switch (benchmark) {
    case 1:  // init_IVC_general.f90:13-16, boundary_ivc_general.f90
        c->dx = 10.0 / ni;
        c->dy = 10.0 / nj;
        c->bc[0] = c->bc[1] = DENSE_BC_EXTRAP;
        c->bc[2] = c->bc[3] = DENSE_BC_PERIODIC;
        return DENSE_OK;
    case 3:  // init_RT_general.f90:9-18, boundary_rt_general.f90
        c->gamma = 5.0 / 3.0;
        c->if_bodyforce = 1;
        c->dx = 0.25 / (ni - 1.0);
        c->dy = 1.0 / (nj - 1.0);
        c->bc[0] = c->bc[1] = DENSE_BC_PERIODIC;
        c->bc[2] = DENSE_BC_DIRICHLET;
        c->bc_state[2][0] = 2.0, c->bc_state[2][1] = 0.0, c->bc_state[2][2] = 0.0, c->bc_state[2][3] = 1.0;
        c->bc[3] = DENSE_BC_DIRICHLET;
        c->bc_state[3][0] = 1.0, c->bc_state[3][1] = 0.0, c->bc_state[3][2] = 0.0, c->bc_state[3][3] = 2.5;
        return DENSE_OK;
    case 5:  // init_doublemach_general.f90:13-16, boundary_doublemach_general.f90
        c->dx = 4.0 / ni;
        c->dy = 1.0 / nj;
        c->bc[0] = DENSE_BC_DIRICHLET;
        c->bc[1] = DENSE_BC_EXTRAP;
        c->bc[2] = DENSE_BC_DMR_WALL;
        c->bc[3] = DENSE_BC_DMR_TOP;
        for (int s = 0; s < 4; s++)
            for (int v = 0; v < 4; v++) c->bc_state[s][v] = post[v];
        return DENSE_OK;
    case 9:  // init_shockvortex_general.f90:13-16, boundary_shockvortex_general.f90
        c->dx = 2.0 / ni;
        c->dy = 1.0 / nj;
        c->bc[0] = c->bc[1] = DENSE_BC_EXTRAP;
        c->bc[2] = c->bc[3] = DENSE_BC_REFLECT;
        return DENSE_OK;
    case 2:  // Sedov blast, init_sedov_general.f90:13-16: reflective left/bottom, outlet right/top
        c->dx = 1.2 / ni;
        c->dy = 1.2 / nj;
        c->bc[DENSE_SIDE_LEFT] = c->bc[DENSE_SIDE_BOTTOM] = DENSE_BC_REFLECT;
        c->bc[DENSE_SIDE_RIGHT] = c->bc[DENSE_SIDE_TOP] = DENSE_BC_EXTRAP;
        return DENSE_OK;
    case 6:  // shear layer, init_ssl_general.f90:12-15: inflow left (boundary_ssl_general.f90:55-85),
             // outlet right, reflective bottom; the top ghost cells keep the state the
             // initialisation gives them (init_ssl_general.f90:62-81; the top loop of
             // boundary_ssl_general.f90:47-52 is empty): a fixed state
        c->dx = 200.0 / ni;
        c->dy = 40.0 / nj;
        c->bc[DENSE_SIDE_LEFT] = DENSE_BC_SSL_INLET;
        c->bc_state[DENSE_SIDE_LEFT][0] = 40.0 / 2.0;
        c->bc[DENSE_SIDE_RIGHT] = DENSE_BC_EXTRAP;
        c->bc[DENSE_SIDE_BOTTOM] = DENSE_BC_REFLECT;
        c->bc[DENSE_SIDE_TOP] = DENSE_BC_DIRICHLET;
        c->bc_state[DENSE_SIDE_TOP][0] = 2.1101, c->bc_state[DENSE_SIDE_TOP][1] = 2.9709;
        c->bc_state[DENSE_SIDE_TOP][2] = -0.1367, c->bc_state[DENSE_SIDE_TOP][3] = (double)0.4754f;
        return DENSE_OK;
    case 4:  // Richtmyer-Meshkov, init_RM_general.f90:12-15: outlet left/right, reflective bottom/top
        c->dx = 4.0 / ni;
        c->dy = 1.0 / nj;
        c->bc[0] = c->bc[1] = DENSE_BC_EXTRAP;
        c->bc[2] = c->bc[3] = DENSE_BC_REFLECT;
        return DENSE_OK;
    case 0:
        return DENSE_OK;
    default:
        return fail(nullptr, DENSE_ERR_UNSUPPORTED, "no such benchmark index %d", benchmark);
    }
}

static StageArgs base_args(const dense_solver* h)
{
    StageArgs a{};
    const dense_cfg& c = h->cfg;
    a.st = h->st;
    a.dt_hist = h->dt_hist;
    a.ni = c.ni;
    a.nj = c.nj;
    a.ld = h->ld;
    a.plane = h->plane;
    a.dx = c.dx;
    a.dy = c.dy;
    a.gamma = c.gamma;
    a.cfl = c.cfl;
    a.delta = c.delta;
    a.dt_const = c.dt;
    a.sqrt3 = std::sqrt(3.0);
    a.if_const_dt = c.if_const_dt;
    a.if_bodyforce = c.if_bodyforce;
    a.tiles_x = h->tiles_x;
    a.ntiles = h->ntiles;
    bc_fill(a.bc, c.bc, c.bc_state, c.gamma);
    a.bc_rows = h->bc_rows;
    return a;
}

// Queue one time step: stages 1..S; the last stage writes U^n in place
// (it reads U^n only at the cell it writes).
static int enqueue_step(dense_solver* h)
{
    StageArgs a = base_args(h);
    const int S = h->cfg.stages;
    for (int s = 1; s <= S; s++) {
        a.stage = s;
        a.last_stage = (s == S);
        a.un = h->U[0];
        a.us = s == 1 ? h->U[0] : h->U[s - 1];
        a.out = (s == S && S > 1) ? h->U[0] : h->U[s];
        HIPCHK(h, launch_stage(h->scheme, h->tile, a, h->stream));
    }
    if (S == 1)
        HIPCHK(h, hipMemcpyAsync(h->U[0], h->U[1], sizeof(double) * 4 * h->plane, hipMemcpyDeviceToDevice, h->stream));
    return DENSE_OK;
}

static void release_graph(dense_solver* h)
{
    if (h->graph) (void)hipGraphExecDestroy(h->graph);
    h->graph = nullptr;
    h->graph_steps = 0;
    for (int r = 0; r < 16; r++) {
        if (h->graph_rem[r]) (void)hipGraphExecDestroy(h->graph_rem[r]);
        h->graph_rem[r] = nullptr;
        h->graph_rem_steps[r] = 0;
    }
}

extern "C" int dense_create(const dense_cfg* cfg, dense_handle* out)
{
    if (!cfg || !out) return fail(nullptr, DENSE_ERR_ARG, "null argument");
    *out = nullptr;
    dense_cfg c = *cfg;
    if (c.benchmark != 0) {
        dense_cfg d = c;
        int rc = dense_problem_defaults(c.benchmark, c.ni, c.nj, &d);
        if (rc) return rc;
        // metrics/BCs from the canned problem; run switches from the caller
        for (int s = 0; s < 4; s++) {
            c.bc[s] = d.bc[s];
            for (int v = 0; v < 4; v++) c.bc_state[s][v] = d.bc_state[s][v];
        }
        if (!(c.dx > 0)) c.dx = d.dx;
        if (!(c.dy > 0)) c.dy = d.dy;
    }
    char why[256];
    int rc = validate_cfg(&c, why, sizeof why);
    if (rc) return fail(nullptr, rc, "%s", why);
    dense_solver* h = new (std::nothrow) dense_solver;
    if (!h) return fail(nullptr, DENSE_ERR_NOMEM, "host allocation failed");
    h->cfg = c;
    h->scheme = scheme_of(c.splitting, c.recon);
    h->ld = (c.ni + 31) / 32 * 32;
    h->plane = (long long)h->ld * c.nj;
    h->tile = resolve_tile(c.tile, c.splitting, (long long)c.ni * c.nj);
    h->ntiles = tile_count(h->tile, c.ni, c.nj, &h->tiles_x);
    auto bail = [&](int code) {
        dense_destroy(h);
        return code;
    };
    if (hipStreamCreateWithFlags(&h->stream, hipStreamNonBlocking) != hipSuccess)
        return bail(fail(nullptr, DENSE_ERR_HIP, "hipStreamCreate failed"));
    for (int k = 0; k < 3; k++)
        if (hipMalloc(&h->U[k], sizeof(double) * 4 * h->plane) != hipSuccess)
            return bail(fail(nullptr, DENSE_ERR_NOMEM, "hipMalloc of state failed"));
    if (hipEventCreate(&h->ev0) != hipSuccess || hipEventCreate(&h->ev1) != hipSuccess)
        return bail(fail(nullptr, DENSE_ERR_HIP, "hipEventCreate failed"));
    if (hipMalloc(&h->st, sizeof(DevState)) != hipSuccess || hipHostMalloc(&h->st_host, sizeof(DevState)) != hipSuccess)
        return bail(fail(nullptr, DENSE_ERR_NOMEM, "state allocation failed"));
    h->hist_cap = 4096;
    if (hipMalloc(&h->dt_hist, sizeof(double) * h->hist_cap) != hipSuccess)
        return bail(fail(nullptr, DENSE_ERR_NOMEM, "dt history allocation failed"));
    for (int k = 0; k < 3; k++)
        if (hipMemsetAsync(h->U[k], 0, sizeof(double) * 4 * h->plane, h->stream) != hipSuccess)
            return bail(fail(nullptr, DENSE_ERR_HIP, "memset failed"));
    DevState z{};
    z.hist_cap = h->hist_cap;
    z.t_end = 0.0;
    if (hipMemcpyAsync(h->st, &z, sizeof z, hipMemcpyHostToDevice, h->stream) != hipSuccess ||
        hipStreamSynchronize(h->stream) != hipSuccess)
        return bail(fail(nullptr, DENSE_ERR_HIP, "state init failed"));
    if (c.bc[DENSE_SIDE_LEFT] == DENSE_BC_SSL_INLET) {
        // per-row terms of the inflow, boundary_ssl_general.f90:58-71, with the
        // cell-centre y of init_ssl_general.f90:20 (host libm, as the reference)
        std::vector<double> rows(2 * (size_t)c.nj);
        for (int j = 0; j < c.nj; j++) {
            const double y = ((double)(j + 1) - 0.5) * c.dy - c.bc_state[DENSE_SIDE_LEFT][0];
            rows[j] = 2.5 + 0.5 * std::tanh(2 * y);
            rows[c.nj + j] = std::exp(-(y * y) / 10.0);
        }
        if (hipMalloc(&h->bc_rows, sizeof(double) * rows.size()) != hipSuccess ||
            hipHostMalloc(&h->bct_host, sizeof(double) * 2) != hipSuccess)
            return bail(fail(nullptr, DENSE_ERR_NOMEM, "inflow table allocation failed"));
        if (hipMemcpy(h->bc_rows, rows.data(), sizeof(double) * rows.size(), hipMemcpyHostToDevice) != hipSuccess)
            return bail(fail(nullptr, DENSE_ERR_HIP, "inflow table upload failed"));
    }
    *out = h;
    return DENSE_OK;
}

extern "C" int dense_destroy(dense_handle h)
{
    if (!h) return DENSE_OK;
    if (h->stream) (void)hipStreamSynchronize(h->stream);
    release_graph(h);
    for (int k = 0; k < 3; k++)
        if (h->U[k]) (void)hipFree(h->U[k]);
    if (h->st) (void)hipFree(h->st);
    if (h->st_host) (void)hipHostFree(h->st_host);
    if (h->dt_hist) (void)hipFree(h->dt_hist);
    if (h->pad_dev) (void)hipFree(h->pad_dev);
    if (h->theta_dev) (void)hipFree(h->theta_dev);
    if (h->pad_host) (void)hipHostFree(h->pad_host);
    if (h->stage_host) (void)hipHostFree(h->stage_host);
    if (h->ev0) (void)hipEventDestroy(h->ev0);
    if (h->ev1) (void)hipEventDestroy(h->ev1);
    if (h->bc_rows) (void)hipFree(h->bc_rows);
    if (h->bct_host) (void)hipHostFree(h->bct_host);
    if (h->stream) (void)hipStreamDestroy(h->stream);
    delete h;
    return DENSE_OK;
}

static int read_state(dense_solver* h)
{
    HIPCHK(h, hipMemcpyAsync(h->st_host, h->st, sizeof(DevState), hipMemcpyDeviceToHost, h->stream));
    HIPCHK(h, hipStreamSynchronize(h->stream));
    return DENSE_OK;
}

extern "C" int dense_set_state(dense_handle h, const double* cons)
{
    if (!h || !cons) return fail(h, DENSE_ERR_ARG, "null argument");
    const int ni = h->cfg.ni, nj = h->cfg.nj, nip = ni + 2 * BF, njp = nj + 2 * BF;
    if (!h->stage_host && hipHostMalloc(&h->stage_host, sizeof(double) * 4 * h->plane) != hipSuccess)
        return fail(h, DENSE_ERR_NOMEM, "pinned staging allocation failed");
    HIPCHK(h, hipStreamSynchronize(h->stream));
    for (int k = 0; k < 4; k++)
        for (int j = 0; j < nj; j++) {
            const double* src = cons + (size_t)k * nip * njp + (size_t)(j + BF) * nip + BF;
            double* dst = h->stage_host + k * h->plane + (long long)j * h->ld;
            std::memcpy(dst, src, sizeof(double) * ni);
            for (int i = ni; i < h->ld; i++) dst[i] = 0.0;
        }
    HIPCHK(h, hipMemcpyAsync(h->U[0], h->stage_host, sizeof(double) * 4 * h->plane, hipMemcpyHostToDevice, h->stream));
    h->stepped = false;
    h->stage_in_valid = false;
    // fresh reduction ring, then the wave speeds of the uploaded state
    int rc = read_state(h);
    if (rc) return rc;
    DevState z = *h->st_host;
    std::memset(z.alpha, 0, sizeof z.alpha);
    z.seq = 0;
    z.blocks_done.v = 0;
    z.error = 0;
    z.active = 0;
    HIPCHK(h, hipMemcpyAsync(h->st, &z, sizeof z, hipMemcpyHostToDevice, h->stream));
    StageArgs a = base_args(h);
    a.us = h->U[0];
    HIPCHK(h, launch_max(a, h->stream));
    HIPCHK(h, hipStreamSynchronize(h->stream));
    return DENSE_OK;
}

// Padded cons (and prim, entropy included) into the caller's arrays: interior
// from U^n, ghost layers from the boundary rules applied to `gsrc` at time tg.
static int padded_out(dense_solver* h, const double* gsrc, double tg, double* cons, double* prim)
{
    const int ni = h->cfg.ni, nj = h->cfg.nj;
    const size_t np = (size_t)(ni + 2 * BF) * (nj + 2 * BF);
    if (!h->pad_dev && hipMalloc(&h->pad_dev, sizeof(double) * 9 * np) != hipSuccess)
        return fail(h, DENSE_ERR_NOMEM, "padded buffer allocation failed");
    if (!h->pad_host && hipHostMalloc(&h->pad_host, sizeof(double) * 9 * np) != hipSuccess)
        return fail(h, DENSE_ERR_NOMEM, "pinned padded buffer allocation failed");
    if (h->bc_rows) {
        // inflow ghost cells at tg
        ssl_clock_terms(tg, h->bct_host);
        HIPCHK(h, hipMemcpyAsync(&h->st->bc_t[0], h->bct_host, sizeof(double) * 2, hipMemcpyHostToDevice, h->stream));
    }
    StageArgs a = base_args(h);
    a.us = h->U[0];
    HIPCHK(h, launch_fill_padded(a, gsrc, tg, h->pad_dev, h->pad_dev + 4 * np, h->stream));
    HIPCHK(h, hipMemcpyAsync(h->pad_host, h->pad_dev, sizeof(double) * 9 * np, hipMemcpyDeviceToHost, h->stream));
    HIPCHK(h, hipStreamSynchronize(h->stream));
    std::memcpy(cons, h->pad_host, sizeof(double) * 4 * np);
    if (prim) {
        std::memcpy(prim, h->pad_host + 4 * np, sizeof(double) * 5 * np);
        // entropy s = ln p - ln rho on interior cells (get_primary.f90:67)
        const int nip = ni + 2 * BF;
        double* s = prim + 5 * np;
        std::memset(s, 0, sizeof(double) * np);
        for (int j = BF; j < nj + BF; j++)
            for (int i = BF; i < ni + BF; i++) {
                const size_t o = (size_t)j * nip + i;
                s[o] = std::log(prim[3 * np + o]) - std::log(prim[o]);
            }
    }
    return DENSE_OK;
}

extern "C" int dense_get_state(dense_handle h, double* cons, double* prim)
{
    if (!h || !cons) return fail(h, DENSE_ERR_ARG, "null argument");
    int rc = read_state(h);
    if (rc) return rc;
    return padded_out(h, h->U[0], h->st_host->t, cons, prim);
}

// cons as the reference's array holds it at a save: interior = the current
// state; ghost layers = the last boundary pass, i.e. the boundary rules on the
// last stage's input at t^n (general_main.f90:26-31, written by
// write_date_binary at :43-44).
extern "C" int dense_get_dump(dense_handle h, double* cons)
{
    if (!h || !cons) return fail(h, DENSE_ERR_ARG, "null argument");
    int rc = read_state(h);
    if (rc) return rc;
    if (!h->stage_in_valid) return padded_out(h, h->U[0], h->st_host->t, cons, nullptr);
    if (h->cfg.stages < 2)
        return fail(h, DENSE_ERR_UNSUPPORTED, "one-stage steps keep no copy of the last stage's input");
    return padded_out(h, h->U[h->cfg.stages - 1], h->st_host->t_prev, cons, nullptr);
}

extern "C" int dense_step(dense_handle h, double current_time, double* dt)
{
    if (!h) return fail(h, DENSE_ERR_ARG, "null handle");
    double tc[2] = {0.0, 0.0};
    if (h->bc_rows) ssl_clock_terms(current_time, tc);
    HIPCHK(h, launch_set_clock(h->st, current_time, INFINITY, 0, 1, h->stream, tc[0], tc[1]));
    int rc = enqueue_step(h);
    if (rc) return rc;
    rc = read_state(h);
    if (rc) return rc;
    h->stepped = true;
    h->stage_in_valid = true;
    if (dt) *dt = h->st_host->dt;
    if (h->st_host->error) return fail(h, DENSE_ERR_NONFINITE, "non-finite wave speed after step");
    return DENSE_OK;
}

// Capture nsteps steps into *exec (replacing what it held).
static int build_graph(dense_solver* h, int nsteps, hipGraphExec_t* exec, int* exec_steps)
{
    if (*exec) (void)hipGraphExecDestroy(*exec);
    *exec = nullptr;
    *exec_steps = 0;
    hipGraph_t g = nullptr;
    HIPCHK(h, hipStreamBeginCapture(h->stream, hipStreamCaptureModeThreadLocal));
    for (int i = 0; i < nsteps; i++) {
        int rc = enqueue_step(h);
        if (rc) {
            (void)hipStreamEndCapture(h->stream, &g);
            if (g) (void)hipGraphDestroy(g);
            return rc;
        }
    }
    HIPCHK(h, hipStreamEndCapture(h->stream, &g));
    hipError_t e = hipGraphInstantiate(exec, g, nullptr, nullptr, 0);
    (void)hipGraphDestroy(g);
    if (e != hipSuccess) {
        *exec = nullptr;
        return fail(h, DENSE_ERR_HIP, "hipGraphInstantiate: %s", hipGetErrorString(e));
    }
    *exec_steps = nsteps;
    return DENSE_OK;
}

// The device-resident dt history holds at least n entries (the captured
// graphs carry its address, so they are rebuilt after it moves).
static int ensure_hist(dense_solver* h, int n)
{
    if (n <= h->hist_cap) return DENSE_OK;
    HIPCHK(h, hipStreamSynchronize(h->stream));
    double* p = nullptr;
    if (hipMalloc(&p, sizeof(double) * (size_t)n) != hipSuccess)
        return fail(h, DENSE_ERR_NOMEM, "dt history of %d steps: allocation failed", n);
    (void)hipFree(h->dt_hist);
    h->dt_hist = p;
    h->hist_cap = n;
    HIPCHK(h, hipMemcpyAsync(&h->st->hist_cap, &h->hist_cap, sizeof(int), hipMemcpyHostToDevice, h->stream));
    HIPCHK(h, hipStreamSynchronize(h->stream));
    release_graph(h);
    return DENSE_OK;
}

constexpr int GRAPH_CHUNK = 16;
static_assert(GRAPH_CHUNK <= 16, "one cached remainder graph per count below the chunk");

// The loop general_main.f90:23-40 on the device: up to max_steps steps while
// t < total_time, step < total_step and !(elasp >= dt_save), the clock, step
// count and save clock advanced by the last stage of every step.
static int run_steps(dense_solver* h, int max_steps, double total_time, int total_step, double dt_save,
                     double* elasp, double* t, int* step, double* dt_hist, int* steps_taken)
{
    if (steps_taken) *steps_taken = 0;
    if (max_steps == 0) return DENSE_OK;
    double el = elasp ? *elasp : 0.0;
    if (h->bc_rows) {
        // the inflow needs the host's cos of t^n: one synchronised step at a time,
        // the clock advanced as the device advances it (t + dt)
        double tt = *t;
        int st = *step, taken = 0;
        HIPCHK(h, hipEventRecord(h->ev0, h->stream));
        while (taken < max_steps && tt < total_time && st < total_step && !(taken > 0 && el >= dt_save)) {
            double dt = 0.0;
            int rc = dense_step(h, tt, &dt);
            if (rc) return rc;
            if (dt_hist) dt_hist[taken] = dt;
            tt = tt + dt;
            el = el + dt;
            st += 1;
            taken += 1;
        }
        HIPCHK(h, hipEventRecord(h->ev1, h->stream));
        HIPCHK(h, hipStreamSynchronize(h->stream));
        // leave the device clock where the step loop left it
        HIPCHK(h, launch_set_clock(h->st, tt, INFINITY, st, st, h->stream));
        HIPCHK(h, hipStreamSynchronize(h->stream));
        *t = tt;
        *step = st;
        if (elasp) *elasp = el;
        if (steps_taken) *steps_taken = taken;
        float ms = 0.f;
        HIPCHK(h, hipEventElapsedTime(&ms, h->ev0, h->ev1));
        h->last_run_ms = ms;
        h->last_run_launches = taken * h->cfg.stages;
        return DENSE_OK;
    }
    if (dt_hist) {
        int rc = ensure_hist(h, max_steps);
        if (rc) return rc;
    }
    // whole chunks, then one graph of the remainder: no replayed step is idle
    // unless the clock or the save clock stops the run early
    const int full = max_steps / GRAPH_CHUNK, rem = max_steps % GRAPH_CHUNK;
    if (full > 0 && h->graph_steps != GRAPH_CHUNK) {
        int rc = build_graph(h, GRAPH_CHUNK, &h->graph, &h->graph_steps);
        if (rc) return rc;
    }
    if (rem > 0 && h->graph_rem_steps[rem] != rem) {
        int rc = build_graph(h, rem, &h->graph_rem[rem], &h->graph_rem_steps[rem]);
        if (rc) return rc;
    }
    // the device stops itself once the clock passes total_time/total_step/dt_save
    const long long want = (long long)*step + max_steps;  // no int overflow near INT_MAX
    const int step_end = want < total_step ? (int)want : total_step;
    HIPCHK(h, launch_set_clock(h->st, *t, total_time, *step, step_end, h->stream, 0.0, 0.0, el, dt_save));
    HIPCHK(h, hipEventRecord(h->ev0, h->stream));
    for (int i = 0; i < full; i++) HIPCHK(h, hipGraphLaunch(h->graph, h->stream));
    if (rem > 0) HIPCHK(h, hipGraphLaunch(h->graph_rem[rem], h->stream));
    HIPCHK(h, hipEventRecord(h->ev1, h->stream));
    int rc = read_state(h);
    if (rc) return rc;
    const DevState& s = *h->st_host;
    const int taken = s.step - *step;
    if (dt_hist && taken > 0)
        HIPCHK(h, hipMemcpy(dt_hist, h->dt_hist, sizeof(double) * (size_t)taken, hipMemcpyDeviceToHost));
    *t = s.t;
    *step = s.step;
    if (elasp) *elasp = s.elasp;
    if (taken > 0) h->stepped = h->stage_in_valid = true;
    if (steps_taken) *steps_taken = taken;
    float ms = 0.f;
    HIPCHK(h, hipEventElapsedTime(&ms, h->ev0, h->ev1));
    h->last_run_ms = ms;
    h->last_run_launches = taken * h->cfg.stages;
    if (s.error) return fail(h, DENSE_ERR_NONFINITE, "non-finite wave speed at step %d", s.step);
    return DENSE_OK;
}

extern "C" int dense_run(dense_handle h, int max_steps, double total_time, int total_step, double* t, int* step,
                         double* dt_hist, int* steps_taken)
{
    if (!h || !t || !step || max_steps < 0) return fail(h, DENSE_ERR_ARG, "bad arguments");
    return run_steps(h, max_steps, total_time, total_step, INFINITY, nullptr, t, step, dt_hist, steps_taken);
}

extern "C" int dense_run_saving(dense_handle h, int max_steps, double total_time, int total_step, double dt_save,
                                double* elasp, double* t, int* step, double* dt_hist, int* steps_taken)
{
    if (!h || !t || !step || !elasp || max_steps < 0) return fail(h, DENSE_ERR_ARG, "bad arguments");
    return run_steps(h, max_steps, total_time, total_step, dt_save, elasp, t, step, dt_hist, steps_taken);
}

extern "C" int dense_last_run_gpu_ms(dense_handle h, double* ms, int* stage_launches)
{
    if (!h || !ms) return fail(h, DENSE_ERR_ARG, "null argument");
    *ms = h->last_run_ms;
    if (stage_launches) *stage_launches = h->last_run_launches;
    return DENSE_OK;
}

extern "C" int dense_max_eig(dense_handle h, double out[2])
{
    if (!h || !out) return fail(h, DENSE_ERR_ARG, "null argument");
    int rc = read_state(h);
    if (rc) return rc;
    const DevState& s = *h->st_host;
    const int slot = s.seq & 3;
    std::memcpy(&out[0], &s.alpha[slot][0], sizeof(double));
    std::memcpy(&out[1], &s.alpha[slot][1], sizeof(double));
    return DENSE_OK;
}

extern "C" int dense_update_params(dense_handle h, const dense_cfg* cfg)
{
    if (!h || !cfg) return fail(h, DENSE_ERR_ARG, "null argument");
    dense_cfg c = h->cfg;
    c.gamma = cfg->gamma;
    c.cfl = cfg->cfl;
    c.delta = cfg->delta;
    c.dt = cfg->dt;
    c.if_const_dt = cfg->if_const_dt;
    c.if_bodyforce = cfg->if_bodyforce;
    c.splitting = cfg->splitting;
    c.recon = cfg->recon;
    char why[256];
    int rc = validate_cfg(&c, why, sizeof why);
    if (rc) return fail(h, rc, "%s", why);
    HIPCHK(h, hipStreamSynchronize(h->stream));
    h->cfg = c;
    h->scheme = scheme_of(c.splitting, c.recon);
    h->tile = resolve_tile(c.tile, c.splitting, (long long)c.ni * c.nj);
    h->ntiles = tile_count(h->tile, c.ni, c.nj, &h->tiles_x);
    release_graph(h);  // kernel arguments changed
    h->stepped = false;  // the wave-speed ring restarts below
    // the wave speeds of the current state depend on gamma
    int r2 = read_state(h);
    if (r2) return r2;
    StageArgs a = base_args(h);
    a.us = h->U[0];
    DevState z = *h->st_host;
    std::memset(z.alpha, 0, sizeof z.alpha);
    z.seq = 0;
    z.blocks_done.v = 0;
    HIPCHK(h, hipMemcpyAsync(h->st, &z, sizeof z, hipMemcpyHostToDevice, h->stream));
    HIPCHK(h, launch_max(a, h->stream));
    HIPCHK(h, hipStreamSynchronize(h->stream));
    return DENSE_OK;
}

// Time every tile shape on trial_steps steps of the current state (state and
// clock restored after each trial) and keep the fastest as an explicit
// cfg.tile.  Every shape gives the same bits; which is fastest depends on how
// the grid's tile count fills the resident workgroups, not only on its size.
extern "C" int dense_autotune(dense_handle h, int trial_steps, int* tile, double* best_ms)
{
    if (!h || trial_steps < 1) return fail(h, DENSE_ERR_ARG, "bad arguments");
    HIPCHK(h, hipStreamSynchronize(h->stream));
    int rc = read_state(h);
    if (rc) return rc;
    const DevState st0 = *h->st_host;
    const size_t bytes = sizeof(double) * 4 * h->plane;
    double* save = nullptr;
    HIPCHK(h, hipMalloc(&save, bytes));
    auto restore = [&]() -> int {
        HIPCHK(h, hipMemcpyAsync(h->U[0], save, bytes, hipMemcpyDeviceToDevice, h->stream));
        HIPCHK(h, hipMemcpyAsync(h->st, &st0, sizeof st0, hipMemcpyHostToDevice, h->stream));
        HIPCHK(h, hipStreamSynchronize(h->stream));
        return DENSE_OK;
    };
    auto use_tile = [&](int t) {
        h->cfg.tile = t;
        h->tile = resolve_tile(t, h->cfg.splitting, (long long)h->cfg.ni * h->cfg.nj);
        h->ntiles = tile_count(h->tile, h->cfg.ni, h->cfg.nj, &h->tiles_x);
        release_graph(h);
    };
    const int keep = h->cfg.tile;
    double best = 1e300;
    int bt = 0;
    if (hipMemcpyAsync(save, h->U[0], bytes, hipMemcpyDeviceToDevice, h->stream) != hipSuccess ||
        hipStreamSynchronize(h->stream) != hipSuccess)
        rc = fail(h, DENSE_ERR_HIP, "state snapshot failed");
    for (int t : {4, 3, 1, 2, 5}) {  // 16x15, 32x15, 32x7, 8x7, 16x15 with a thread per face
        if (rc) break;
        use_tile(t);
        for (int pass = 0; pass < 2 && !rc; pass++) {  // the first pass builds the graph
            double tt = st0.t;
            int ss = st0.step, taken = 0;
            if (!(rc = restore())) rc = dense_run(h, trial_steps, INFINITY, INT_MAX, &tt, &ss, nullptr, &taken);
        }
        if (!rc && h->last_run_ms < best) {
            best = h->last_run_ms;
            bt = t;
        }
    }
    use_tile(rc ? keep : bt);
    const int rr = restore();
    // the trials overwrote the stage buffers: theta and the dump ghosts of the
    // last step are gone until the next step
    h->stepped = false;
    h->stage_in_valid = false;
    (void)hipFree(save);
    if (rc) return rc;
    if (rr) return rr;
    if (tile) *tile = bt;
    if (best_ms) *best_ms = best;
    return DENSE_OK;
}

// AdaWENO's indicator theta at every face, as the reference's theta arrays hold
// it at output time: from the last flux evaluation, i.e. the input state of the
// last RK stage (U^(2) of a 3-stage step, U^(1) of a 2-stage one) with its
// boundary values at t^n and its global wave speeds.  Fills theta_xip(ni+1,nj),
// theta_xin(ni+1,nj), theta_etap(nj+1,ni), theta_etan(nj+1,ni), Fortran order.
extern "C" int dense_get_theta(dense_handle h, double* theta)
{
    if (!h || !theta) return fail(h, DENSE_ERR_ARG, "null argument");
    if (h->cfg.splitting != 105)
        return fail(h, DENSE_ERR_UNSUPPORTED, "theta is the indicator of AdaWENO (splitting 105)");
    if (h->cfg.stages < 2)
        return fail(h, DENSE_ERR_UNSUPPORTED, "theta needs the input state of the last stage (stages 2 or 3)");
    if (!h->stepped) return fail(h, DENSE_ERR_ARG, "no step since the state or the parameters were set");
    const int ni = h->cfg.ni, nj = h->cfg.nj;
    const size_t n = 2 * ((size_t)(ni + 1) * nj + (size_t)(nj + 1) * ni);
    if (!h->theta_dev && hipMalloc(&h->theta_dev, sizeof(double) * n) != hipSuccess)
        return fail(h, DENSE_ERR_NOMEM, "theta buffer allocation failed");
    int rc = read_state(h);
    if (rc) return rc;
    const DevState s = *h->st_host;
    const int slot = (s.seq - 1) & 3;  // read by the last stage launch
    double ai, aj;
    std::memcpy(&ai, &s.alpha[slot][0], sizeof(double));
    std::memcpy(&aj, &s.alpha[slot][1], sizeof(double));
    if (h->bc_rows) {
        // the inflow ghost cells at t^n (a get_state since may have set them for t^(n+1))
        ssl_clock_terms(s.t_prev, h->bct_host);
        HIPCHK(h, hipMemcpyAsync(&h->st->bc_t[0], h->bct_host, sizeof(double) * 2, hipMemcpyHostToDevice, h->stream));
    }
    StageArgs a = base_args(h);
    a.us = h->U[h->cfg.stages - 1];
    HIPCHK(h, launch_theta(a, ai, aj, s.t_prev, h->theta_dev, h->stream));
    HIPCHK(h, hipMemcpyAsync(theta, h->theta_dev, sizeof(double) * n, hipMemcpyDeviceToHost, h->stream));
    HIPCHK(h, hipStreamSynchronize(h->stream));
    return DENSE_OK;
}

extern "C" int dense_sync(dense_handle h)
{
    if (!h) return fail(h, DENSE_ERR_ARG, "null handle");
    HIPCHK(h, hipStreamSynchronize(h->stream));
    return DENSE_OK;
}

extern "C" const char* dense_last_error(dense_handle h) { return h ? h->err : g_err; }

// ---------------------------------------------------------------------------
// Allocation-free stepping (dense_run_noalloc).  The scratch memory of a run
// -- the device clock and launch counters (DevState) and the two stage
// buffers U^(1), U^(2) -- lives in one file-scope device array instead of
// memory the library allocates.  Host code never passes `scratch` itself to a
// kernel (the name is a host-side stand-in): its device address is looked up
// once per device with hipGetSymbolAddress and kept in g_scratch[device].
// Every call on a device shares that device's array and leaves its contents
// behind; each call clears the clock and counters it needs with a kernel, and
// calls that overlap on one device race.
namespace {
constexpr long long NOALLOC_MAX_CELLS = 1ll << 20;  // ld * nj per stage buffer
constexpr long long NOALLOC_STATE_DOUBLES = 512;    // 4 KB for the DevState
static_assert(sizeof(DevState) <= NOALLOC_STATE_DOUBLES * sizeof(double), "DevState fits its slot");
constexpr long long NOALLOC_DOUBLES = NOALLOC_STATE_DOUBLES + 2 * 4 * NOALLOC_MAX_CELLS;  // 64 MiB + 4 KB
constexpr int NOALLOC_MAX_DEVICES = 64;
}  // namespace

static __device__ __align__(256) double scratch[NOALLOC_DOUBLES];

namespace {
std::atomic<double*> g_scratch[NOALLOC_MAX_DEVICES];

// Device address of the current device's scratch array (looked up on the
// first call on that device).
hipError_t scratch_ptr(double** out)
{
    int dev = 0;
    hipError_t e = hipGetDevice(&dev);
    if (e != hipSuccess) return e;
    if (dev < 0 || dev >= NOALLOC_MAX_DEVICES) return hipErrorInvalidDevice;
    double* p = g_scratch[dev].load(std::memory_order_acquire);
    if (!p) {
        e = hipGetSymbolAddress((void**)&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess) return e;
        g_scratch[dev].store(p, std::memory_order_release);
    }
    *out = p;
    return hipSuccess;
}

// A fresh DevState: reduction ring, sequence number, launch counters, error
// flag and history count zero (a kernel, not a memset, so that a caller may
// capture the call in a graph).
__global__ void noalloc_reset_kernel(DevState* st)
{
    for (int s = 0; s < 4; s++) st->alpha[s][0] = st->alpha[s][1] = 0ull;
    for (int x = 0; x < 8; x++) st->qnext[x].v = 0u;
    st->blocks_done.v = 0u;
    st->seq = 0;
    st->error = 0;
    st->active = 0;
    st->hist_n = 0;
    st->hist_cap = 0;
    st->dt = 0.0;
    st->t_prev = 0.0;
}
}  // namespace

extern "C" int dense_run_noalloc(const dense_cfg* cfg, double* u, int ld, int nsteps, double t0, void* stream,
                                 double* t_out, double* dt_out)
{
    if (!cfg || !u || nsteps < 0) return fail(nullptr, DENSE_ERR_ARG, "bad arguments");
    dense_cfg c = *cfg;
    if (c.benchmark != 0) {
        dense_cfg d = c;
        int rc = dense_problem_defaults(c.benchmark, c.ni, c.nj, &d);
        if (rc) return rc;
        for (int s = 0; s < 4; s++) {
            c.bc[s] = d.bc[s];
            for (int v = 0; v < 4; v++) c.bc_state[s][v] = d.bc_state[s][v];
        }
        if (!(c.dx > 0)) c.dx = d.dx;
        if (!(c.dy > 0)) c.dy = d.dy;
    }
    char why[256];
    int rc = validate_cfg(&c, why, sizeof why);
    if (rc) return fail(nullptr, rc, "%s", why);
    if (c.bc[DENSE_SIDE_LEFT] == DENSE_BC_SSL_INLET)
        return fail(nullptr, DENSE_ERR_UNSUPPORTED, "the shear-layer inflow steps from the host (use dense_run)");
    if (ld < c.ni) return fail(nullptr, DENSE_ERR_ARG, "row pitch %d below ni = %d", ld, c.ni);
    const long long plane = (long long)ld * c.nj;
    if (plane > NOALLOC_MAX_CELLS)
        return fail(nullptr, DENSE_ERR_ARG,
                    "%lld cells per plane (ld %d x nj %d) exceed the allocation-free limit of %lld; use dense_run",
                    plane, ld, c.nj, NOALLOC_MAX_CELLS);
    if (nsteps == 0) {
        if (t_out) *t_out = t0;
        if (dt_out) *dt_out = 0.0;
        return DENSE_OK;
    }
    double* base = nullptr;
    hipError_t e = scratch_ptr(&base);
    if (e != hipSuccess) return fail(nullptr, DENSE_ERR_HIP, "scratch address: %s", hipGetErrorString(e));
    DevState* st = reinterpret_cast<DevState*>(base);
    double* U1 = base + NOALLOC_STATE_DOUBLES;
    double* U2 = U1 + 4 * plane;
    hipStream_t s = static_cast<hipStream_t>(stream);

    StageArgs a{};
    a.st = st;
    a.dt_hist = nullptr;
    a.ni = c.ni;
    a.nj = c.nj;
    a.ld = ld;
    a.plane = plane;
    a.dx = c.dx;
    a.dy = c.dy;
    a.gamma = c.gamma;
    a.cfl = c.cfl;
    a.delta = c.delta;
    a.dt_const = c.dt;
    a.sqrt3 = std::sqrt(3.0);
    a.if_const_dt = c.if_const_dt;
    a.if_bodyforce = c.if_bodyforce;
    const int tile = resolve_tile(c.tile, c.splitting, (long long)c.ni * c.nj);
    a.ntiles = tile_count(tile, c.ni, c.nj, &a.tiles_x);
    bc_fill(a.bc, c.bc, c.bc_state, c.gamma);
    a.bc_rows = nullptr;
    const int scheme = scheme_of(c.splitting, c.recon);

    hipLaunchKernelGGL(noalloc_reset_kernel, dim3(1), dim3(1), 0, s, st);
    HIPCHK(nullptr, hipGetLastError());
    a.us = u;
    HIPCHK(nullptr, launch_max(a, s));  // wave speeds of the input state
    HIPCHK(nullptr, launch_set_clock(st, t0, INFINITY, 0, nsteps, s));
    double* Us[3] = {u, U1, U2};
    const int S = c.stages;
    for (int n = 0; n < nsteps; n++) {
        for (int k = 1; k <= S; k++) {
            a.stage = k;
            a.last_stage = (k == S);
            a.un = u;
            a.us = k == 1 ? u : Us[k - 1];
            a.out = (k == S && S > 1) ? u : Us[k];
            HIPCHK(nullptr, launch_stage(scheme, tile, a, s));
        }
        if (S == 1) HIPCHK(nullptr, hipMemcpyAsync(u, U1, sizeof(double) * 4 * plane, hipMemcpyDeviceToDevice, s));
    }
    if (!t_out && !dt_out) return DENSE_OK;  // enqueued only
    DevState h{};
    HIPCHK(nullptr, hipMemcpyAsync(&h, st, sizeof h, hipMemcpyDeviceToHost, s));
    HIPCHK(nullptr, hipStreamSynchronize(s));
    if (t_out) *t_out = h.t;
    if (dt_out) *dt_out = h.dt;
    if (h.error) return fail(nullptr, DENSE_ERR_NONFINITE, "non-finite wave speed by step %d", h.step);
    return DENSE_OK;
}
