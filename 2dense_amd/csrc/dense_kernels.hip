// Fused RK-stage kernel of the explicit update (one block) and its helpers
// (gfx950).  The per-tile work lives in dense_tile.h.
#include <atomic>

#include "dense_tile.h"

namespace dense {

// Persistent RK-stage kernel: the grid holds only as many workgroups as are
// resident at once; each loops over tiles pulled from its XCD's queue, so the
// per-launch work (clock/state reads, the wave-speed reduction and its
// atomics) is paid once per workgroup rather than once per tile.
template <int SCHEME, class T>
__global__ void __launch_bounds__(T::NT) __attribute__((amdgpu_waves_per_eu(min_waves<SCHEME, T>())))
stage_kernel(StageArgs a)
{
    __shared__ TileLds<SCHEME, T> L;
    __shared__ int s_tile;
    DevState* st = a.st;
    // Every word this launch reads from the device state is loaded at once (the
    // whole wave-speed ring beside seq): one memory round trip, not three.
    unsigned long long ring[8];
#pragma unroll
    for (int k = 0; k < 8; k++) ring[k] = (&st->alpha[0][0])[k];
    const int active = st->active, seq = st->seq, step = st->step;
    const double t = st->t, dt_prev = st->dt;
    if (!active) return;
    unsigned long long bi = ring[0], bj = ring[1];
#pragma unroll
    for (int k = 1; k < 4; k++)
        if ((seq & 3) == k) bi = ring[2 * k], bj = ring[2 * k + 1];
    const double ai = __longlong_as_double((long long)bi);
    const double aj = __longlong_as_double((long long)bj);
    double dt;
    if (a.stage == 1) {
        dt = a.dt_sched ? a.dt_sched[step] : cfl_dt(a, ai, aj);
        if (blockIdx.x == 0 && threadIdx.x == 0) st->dt = dt;
    } else {
        dt = dt_prev;  // set by workgroup 0 of this step's first stage
    }
    const int xq = blockIdx.x & 7;
    int lo, hi;
    queue_range(xq, a.ntiles, lo, hi);
    // The first tile of a workgroup is its rank among the workgroups of its
    // XCD (no atomic before the first load); the queue hands out the rest.
    int tile = lo + (int)(blockIdx.x >> 3);
    const int rest = lo + (int)((gridDim.x + 7 - xq) >> 3);
    unsigned long long ei = 0, ej = 0;
    while (tile < hi) {
        run_tile<SCHEME, T>(a, tile, ai, aj, dt, t, L, ei, ej,
                            [&] { s_tile = rest + (int)atomicAdd(&st->qnext[xq].v, 1u); });
        tile = s_tile;
    }
    block_max_to_slot<T::NT>(ei, ej, st->alpha[(seq + 1) & 3]);
    launch_epilogue(st, seq, a.last_stage, t, dt, a.dt_hist);
}

// max|u|+c, max|v|+c of a state (get_primary.f90:69-73), used after a new state is uploaded.
__global__ void __launch_bounds__(256) max_kernel(StageArgs a)
{
    DevState* st = a.st;
    const int seq = st->seq;
    unsigned long long ei = 0, ej = 0;
    const long long n = (long long)a.ni * a.nj;
    for (long long idx = (long long)blockIdx.x * 256 + threadIdx.x; idx < n; idx += (long long)gridDim.x * 256) {
        const int gy = (int)(idx / a.ni), gx = (int)(idx - (long long)gy * a.ni);
        double c[4], pr[4];
        load_interior(a, a.us, gx, gy, c, pr);
        const unsigned long long bi = dbits(fabs(pr[0]) + pr[3]), bj = dbits(fabs(pr[1]) + pr[3]);
        ei = bi > ei ? bi : ei;
        ej = bj > ej ? bj : ej;
    }
    block_max_to_slot<256>(ei, ej, st->alpha[(seq + 1) & 3]);
    launch_epilogue(st, seq, false, 0.0, 0.0, nullptr);
}

// AdaWENO's indicator theta at the faces of one sweep direction: the WENO-Z
// smoothness measure of rho + F+-_n over each face's two 5-point stencils
// (get_theta_xi_p/_n and get_theta_eta_p/_n, scheme_weno5th_SC.f90:1-64 and
// the eta analogues, on the indicator of flux_splitting_SC.f90:58-59, :89-90),
// the same values the stage kernel's AdaWENO faces branch on.  Ghost cells from
// the boundary rules at time t.  Output as the reference's theta arrays, first
// index fastest: DIR 0 theta_xi(ni+1, nj), DIR 1 theta_eta(nj+1, ni); + then -.
template <int DIR>
__global__ void __launch_bounds__(256) theta_kernel(StageArgs a, double alpha, double t, double* thp, double* thn)
{
    constexpr int NRM = DIR == 0 ? 1 : 2;
    const int nf = DIR == 0 ? a.ni + 1 : a.nj + 1, nl = DIR == 0 ? a.nj : a.ni;
    const long long idx = (long long)blockIdx.x * 256 + threadIdx.x;
    if (idx >= (long long)nf * nl) return;
    const int f = (int)(idx % nf), l = (int)(idx / nf);
    double rho[6], mn[6], un[6], p[6];
#pragma unroll
    for (int m = 0; m < 6; m++) {
        const int g = f - 3 + m;  // cell along the sweep: f-3 .. f+2
        const int gx = DIR == 0 ? g : l, gy = DIR == 0 ? l : g;
        double c[4], pr[4];
        if ((unsigned)g < (unsigned)(DIR == 0 ? a.ni : a.nj)) load_interior(a, a.us, gx, gy, c, pr);
        else ghost_cell(a, a.us, gx, gy, t, c, pr);
        rho[m] = c[0];
        mn[m] = c[NRM];
        un[m] = pr[DIR];
        p[m] = pr[2];
    }
    double lf[5], w0, w1, w2, th;
#pragma unroll
    for (int i = 0; i < 5; i++) lf[i] = rho[i] + 0.5 * ((un[i] + alpha) * mn[i] + p[i]);
    weno5z_omega(lf, w0, w1, w2, &th);
    thp[idx] = th;
#pragma unroll
    for (int i = 0; i < 5; i++) lf[i] = rho[5 - i] + 0.5 * ((un[5 - i] - alpha) * mn[5 - i] + p[5 - i]);
    weno5z_omega(lf, w0, w1, w2, &th);
    thn[idx] = th;
}

__global__ void set_clock_kernel(DevState* st, double t, double t_end, int step, int step_end, double bc_t0,
                                 double bc_t1, double elasp, double elasp_end)
{
    st->bc_t[0] = bc_t0;
    st->bc_t[1] = bc_t1;
    st->t = t;
    st->t_end = t_end;
    st->step = step;
    st->step_end = step_end;
    st->elasp = elasp;
    st->elasp_end = elasp_end;
    st->hist_n = 0;
    // the save clock is tested after each step (general_main.f90:40): at least one step runs
    st->active = (t < t_end) && (step < step_end) && !st->error;
}

// Padded reference-layout arrays cons(ni+8,nj+8,4), prim(ni+8,nj+8,5): the
// interior from a.us, the ghost layers from the boundary rules applied to gsrc
// at time tg (the corners, which no boundary routine writes, zero).
__global__ void __launch_bounds__(256) fill_padded_kernel(StageArgs a, const double* gsrc, double tg, double* cons,
                                                          double* prim)
{
    const int nip = a.ni + 2 * BF, njp = a.nj + 2 * BF;
    const long long np = (long long)nip * njp;
    const long long idx = (long long)blockIdx.x * 256 + threadIdx.x;
    if (idx >= np) return;
    const int J = (int)(idx / nip), I = (int)(idx - (long long)J * nip);
    const int gx = I - BF, gy = J - BF;
    const bool inx = (unsigned)gx < (unsigned)a.ni, iny = (unsigned)gy < (unsigned)a.nj;
    double c[4] = {0, 0, 0, 0}, pr[4] = {0, 0, 0, 0};
    double rho = 0.0;
    if (inx && iny) {
        load_interior(a, a.us, gx, gy, c, pr);
        rho = c[0];
    } else if (inx || iny) {
        ghost_cell(a, gsrc, gx, gy, tg, c, pr);
        rho = c[0];
    }
#pragma unroll
    for (int k = 0; k < 4; k++) cons[k * np + idx] = c[k];
    prim[idx] = rho;
#pragma unroll
    for (int k = 0; k < 4; k++) prim[(k + 1) * np + idx] = pr[k];
}

// ------------------------------------------------------------------ launchers

// Persistent grid: the resident capacity of the device (blocks per CU from the
// occupancy calculator times the CU count), a multiple of 8 so that block
// b sits on XCD b % 8, and no larger than needed for the tiles.
template <int SCHEME, class T>
static hipError_t launch_one(const StageArgs& a, hipStream_t s)
{
    // resident blocks of this kernel on the device, computed once (several
    // handles may launch from different host threads: relaxed atomics, every
    // writer stores the same value)
    static std::atomic<int> cached{0};
    int capacity = cached.load(std::memory_order_relaxed);
    if (capacity == 0) {
        int dev = 0, per_cu = 0, cus = 0;
        hipError_t e = hipGetDevice(&dev);
        if (e == hipSuccess) e = hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev);
        if (e == hipSuccess)
            e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, stage_kernel<SCHEME, T>, T::NT, 0);
        if (e != hipSuccess) return e;
        capacity = (per_cu < 1 ? 1 : per_cu) * (cus < 1 ? 1 : cus);
        cached.store(capacity, std::memory_order_relaxed);
    }
    int grid = (a.ntiles + 7) / 8 * 8;
    if (grid > capacity) grid = capacity / 8 * 8;
    if (grid < 8) grid = 8;
    hipLaunchKernelGGL((stage_kernel<SCHEME, T>), dim3(grid), dim3(T::NT), 0, s, a);
    return hipGetLastError();
}

template <class T>
static hipError_t launch_scheme(int scheme, const StageArgs& a, hipStream_t s)
{
    switch (scheme) {
    case S_LLF_UP5: return launch_one<S_LLF_UP5, T>(a, s);
    case S_LLF_JS: return launch_one<S_LLF_JS, T>(a, s);
    case S_LLF_Z: return launch_one<S_LLF_Z, T>(a, s);
    case S_GLF_UP5: return launch_one<S_GLF_UP5, T>(a, s);
    case S_GLF_JS: return launch_one<S_GLF_JS, T>(a, s);
    case S_GLF_Z: return launch_one<S_GLF_Z, T>(a, s);
    case S_SW_UP5: return launch_one<S_SW_UP5, T>(a, s);
    case S_SW_JS: return launch_one<S_SW_JS, T>(a, s);
    case S_SW_Z: return launch_one<S_SW_Z, T>(a, s);
    case S_CH: return launch_one<S_CH, T>(a, s);
    case S_SC: return launch_one<S_SC, T>(a, s);
    default: return hipErrorInvalidValue;
    }
}

int tile_count(int tile, int ni, int nj, int* tiles_x)
{
    int tx, ty, nt;
    tile_shape(tile, tx, ty, nt);
    const int nx = (ni + tx - 1) / tx, ny = (nj + ty - 1) / ty;
    if (tiles_x) *tiles_x = nx;
    return nx * ny;
}

hipError_t launch_stage(int scheme, int tile, const StageArgs& a, hipStream_t s)
{
    switch (tile) {
#define X(i, a_, b_, c_) \
    case i: return launch_scheme<Tile<a_, b_, c_>>(scheme, a, s);
        DENSE_TILE_LIST(X)
#undef X
    default: return hipErrorInvalidValue;
    }
}

hipError_t launch_theta(const StageArgs& a, double ai, double aj, double t, double* theta, hipStream_t s)
{
    const long long nx = (long long)(a.ni + 1) * a.nj, ny = (long long)(a.nj + 1) * a.ni;
    hipLaunchKernelGGL(theta_kernel<0>, dim3((unsigned)((nx + 255) / 256)), dim3(256), 0, s, a, ai, t, theta,
                       theta + nx);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(theta_kernel<1>, dim3((unsigned)((ny + 255) / 256)), dim3(256), 0, s, a, aj, t,
                       theta + 2 * nx, theta + 2 * nx + ny);
    return hipGetLastError();
}

hipError_t launch_max(const StageArgs& a, hipStream_t s)
{
    long long n = (long long)a.ni * a.nj;
    int blocks = (int)((n + 255) / 256);
    if (blocks > 2048) blocks = 2048;
    if (blocks < 1) blocks = 1;
    hipLaunchKernelGGL(max_kernel, dim3(blocks), dim3(256), 0, s, a);
    return hipGetLastError();
}

hipError_t launch_set_clock(DevState* st, double t, double t_end, int step, int step_end, hipStream_t s,
                            double bc_t0, double bc_t1, double elasp, double elasp_end)
{
    hipLaunchKernelGGL(set_clock_kernel, dim3(1), dim3(1), 0, s, st, t, t_end, step, step_end, bc_t0, bc_t1, elasp,
                       elasp_end);
    return hipGetLastError();
}

hipError_t launch_fill_padded(const StageArgs& a, const double* ghost_src, double t_ghost, double* cons, double* prim,
                              hipStream_t s)
{
    long long np = (long long)(a.ni + 2 * BF) * (a.nj + 2 * BF);
    hipLaunchKernelGGL(fill_padded_kernel, dim3((unsigned)((np + 255) / 256)), dim3(256), 0, s, a, ghost_src, t_ghost,
                       cons, prim);
    return hipGetLastError();
}

}  // namespace dense
