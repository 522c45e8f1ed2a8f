// Batched mode: an ensemble of independent blocks of uneven size advanced by
// ONE persistent launch per RK stage.  All tiles of all instances form one
// list; workgroups pull tiles from per-XCD queues, switch instance when a
// tile belongs to another block, and keep per-instance wave-speed maxima and
// clocks, so every instance advances with its own CFL step exactly as it
// would alone (bit for bit: the per-tile work is the single-block code).
#include <algorithm>
#include <atomic>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <new>
#include <numeric>
#include <vector>

#include "dense_tile.h"

namespace dense {

int validate_cfg(const dense_cfg* c, char* why, size_t n);  // dense_capi.hip

template <int SCHEME, class T>
__global__ void __launch_bounds__(T::NT) __attribute__((amdgpu_waves_per_eu(min_waves<SCHEME, T>())))
batch_stage_kernel(BatchArgs b)
{
    __shared__ TileLds<SCHEME, T> L;
    __shared__ StageArgs sa;  // the current instance, as the single-block kernel sees its block
    __shared__ double s_ai, s_aj, s_dt, s_t;
    __shared__ int s_tile, s_last;
    BatchState* bs = b.bst;
    if (!bs->active) return;
    const int seq = bs->seq;
    const int tid = threadIdx.x;
    const int xq = blockIdx.x & 7;
    int lo, hi;
    queue_range(xq, b.ntiles, lo, hi);
    // first tile: the workgroup's rank on its XCD; the queue hands out the rest
    int tile = lo + (int)(blockIdx.x >> 3);
    const int rest = lo + (int)((gridDim.x + 7 - xq) >> 3);
    int cur = -1;
    unsigned long long ei = 0, ej = 0;
    while (tile < hi) {
        const int inst = b.tile_inst[tile];
        if (inst != cur) {
            if (cur >= 0) {
                block_max_to_slot<T::NT>(ei, ej, b.ist[cur].alpha[(seq + 1) & 3]);
                ei = ej = 0;
            }
            if (tid == 0) {
                const BatchInst& I = b.inst[inst];
                BatchInstState& S = b.ist[inst];
                StageArgs a{};
                a.un = I.U[0];
                a.us = b.stage == 1 ? I.U[0] : I.U[b.stage - 1];
                a.out = (b.stage == b.nstages) ? I.U[0] : I.U[b.stage];
                a.ni = I.ni;
                a.nj = I.nj;
                a.ld = I.ld;
                a.plane = I.plane;
                a.dx = I.dx;
                a.dy = I.dy;
                a.gamma = I.gamma;
                a.cfl = I.cfl;
                a.delta = I.delta;
                a.dt_const = I.dt_const;
                a.sqrt3 = b.sqrt3;
                a.if_const_dt = I.if_const_dt;
                a.if_bodyforce = I.if_bodyforce;
                a.stage = b.stage;
                a.last_stage = b.last_stage;
                a.tiles_x = I.tiles_x;
                a.ntiles = I.ntiles;
                a.bc = I.bc;
                sa = a;
                s_ai = __longlong_as_double((long long)S.alpha[seq & 3][0]);
                s_aj = __longlong_as_double((long long)S.alpha[seq & 3][1]);
                s_dt = b.stage == 1 ? cfl_dt(a, s_ai, s_aj) : S.dt;
                s_t = S.t;
            }
            __syncthreads();
            cur = inst;
        }
        const int lt = tile - b.inst[inst].tile_base;
        if (b.stage == 1 && lt == 0 && tid == 0) b.ist[inst].dt = s_dt;  // every tile computes the same dt
        run_tile<SCHEME, T>(sa, lt, s_ai, s_aj, s_dt, s_t, L, ei, ej,
                            [&] { s_tile = rest + (int)atomicAdd(&bs->qnext[xq].v, 1u); });
        tile = s_tile;
    }
    if (cur >= 0) block_max_to_slot<T::NT>(ei, ej, b.ist[cur].alpha[(seq + 1) & 3]);

    // last workgroup: per-instance clocks and slots, then the batch counters
    __syncthreads();
    if (tid == 0) {
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
        wait_vm_writes();
        const unsigned int prev =
            __hip_atomic_fetch_add(&bs->blocks_done.v, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        s_last = (prev == gridDim.x - 1);
    }
    __syncthreads();
    if (!s_last) return;
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");  // the other blocks' maxima (see launch_epilogue)
    const int out = (seq + 1) & 3, nxt = (seq + 2) & 3;
    const unsigned long long inf_bits = 0x7ff0000000000000ull;
    for (int i = tid; i < b.ninst; i += T::NT) {
        BatchInstState& S = b.ist[i];
        __hip_atomic_store(&S.alpha[nxt][0], 0ull, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        __hip_atomic_store(&S.alpha[nxt][1], 0ull, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        const unsigned long long ai = __hip_atomic_load(&S.alpha[out][0], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        const unsigned long long aj = __hip_atomic_load(&S.alpha[out][1], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        if (ai >= inf_bits || aj >= inf_bits) S.error = 1;
        if (b.last_stage) {
            if (b.dt_hist && S.step < b.hist_cap) b.dt_hist[(long long)S.step * b.ninst + i] = S.dt;
            S.t = S.t + S.dt;
            S.step += 1;
        }
    }
    if (tid == 0) {
        __hip_atomic_store(&bs->blocks_done.v, 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
#pragma unroll
        for (int x = 0; x < 8; x++) __hip_atomic_store(&bs->qnext[x].v, 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        bs->seq = seq + 1;
        if (b.last_stage) {
            bs->step += 1;
            bs->active = bs->step < bs->step_end;
        }
    }
}

// Wave speeds of every instance's current state (after uploads).
template <int TX, int TY>
__global__ void __launch_bounds__(256) batch_max_kernel(BatchArgs b)
{
    BatchState* bs = b.bst;
    const int seq = bs->seq;
    const int tile = blockIdx.x;
    const int inst = b.tile_inst[tile];
    const BatchInst& I = b.inst[inst];
    const int lt = tile - I.tile_base;
    const int x0 = (lt % I.tiles_x) * TX, y0 = (lt / I.tiles_x) * TY;
    unsigned long long ei = 0, ej = 0;
    StageArgs a{};
    a.ld = I.ld;
    a.plane = I.plane;
    a.gamma = I.gamma;
    for (int idx = threadIdx.x; idx < TX * TY; idx += 256) {
        const int gx = x0 + idx % TX, gy = y0 + idx / TX;
        if (gx >= I.ni || gy >= I.nj) continue;
        double c[4], pr[4];
        load_interior(a, I.U[0], gx, gy, c, pr);
        const unsigned long long bi = dbits(fabs(pr[0]) + pr[3]), bj = dbits(fabs(pr[1]) + pr[3]);
        ei = bi > ei ? bi : ei;
        ej = bj > ej ? bj : ej;
    }
    block_max_to_slot<256>(ei, ej, b.ist[inst].alpha[(seq + 1) & 3]);
}

__global__ void batch_advance_seq_kernel(BatchState* bs, BatchInstState* ist, int ninst)
{
    const int seq = bs->seq;
    for (int i = threadIdx.x; i < ninst; i += blockDim.x) {
        ist[i].alpha[(seq + 2) & 3][0] = 0ull;
        ist[i].alpha[(seq + 2) & 3][1] = 0ull;
    }
    __syncthreads();
    if (threadIdx.x == 0) bs->seq = seq + 1;
}

__global__ void batch_set_run_kernel(BatchState* bs, int nsteps)
{
    bs->step_end = bs->step + nsteps;
    bs->active = nsteps > 0;
}

template <int SCHEME, class T>
static hipError_t launch_batch_one(int grid_pct, const BatchArgs& b, hipStream_t s)
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
            e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, batch_stage_kernel<SCHEME, T>, T::NT, 0);
        if (e != hipSuccess) return e;
        capacity = (per_cu < 1 ? 1 : per_cu) * (cus < 1 ? 1 : cus);
        cached.store(capacity, std::memory_order_relaxed);
    }
    int grid = (int)((long long)capacity * grid_pct / 100) / 8 * 8;
    const int need = (b.ntiles + 7) / 8 * 8;
    if (grid > need) grid = need;
    if (grid < 8) grid = 8;
    hipLaunchKernelGGL((batch_stage_kernel<SCHEME, T>), dim3(grid), dim3(T::NT), 0, s, b);
    return hipGetLastError();
}

template <class T>
static hipError_t launch_batch_scheme(int scheme, int grid_pct, const BatchArgs& b, hipStream_t s)
{
    if constexpr (T::DIRSPLIT) {
        return hipErrorInvalidValue;  // not a batch schedule (dense_batch_set_schedule takes 0..3)
    } else {
    switch (scheme) {
    case S_LLF_UP5: return launch_batch_one<S_LLF_UP5, T>(grid_pct, b, s);
    case S_LLF_JS: return launch_batch_one<S_LLF_JS, T>(grid_pct, b, s);
    case S_LLF_Z: return launch_batch_one<S_LLF_Z, T>(grid_pct, b, s);
    case S_GLF_UP5: return launch_batch_one<S_GLF_UP5, T>(grid_pct, b, s);
    case S_GLF_JS: return launch_batch_one<S_GLF_JS, T>(grid_pct, b, s);
    case S_GLF_Z: return launch_batch_one<S_GLF_Z, T>(grid_pct, b, s);
    case S_SW_UP5: return launch_batch_one<S_SW_UP5, T>(grid_pct, b, s);
    case S_SW_JS: return launch_batch_one<S_SW_JS, T>(grid_pct, b, s);
    case S_SW_Z: return launch_batch_one<S_SW_Z, T>(grid_pct, b, s);
    case S_CH: return launch_batch_one<S_CH, T>(grid_pct, b, s);
    case S_SC: return launch_batch_one<S_SC, T>(grid_pct, b, s);
    default: return hipErrorInvalidValue;
    }
    }
}

hipError_t launch_batch_stage(int scheme, int tile, int grid_pct, const BatchArgs& b, hipStream_t s)
{
    switch (tile) {
#define X(i, a_, b_, c_) \
    case i: return launch_batch_scheme<Tile<a_, b_, c_>>(scheme, grid_pct, b, s);
        DENSE_TILE_LIST(X)
#undef X
    default: return hipErrorInvalidValue;
    }
}

static hipError_t launch_batch_max_t(int tile, const BatchArgs& b, hipStream_t s)
{
    switch (tile) {
#define X(i, a_, b_, c_) \
    case i: hipLaunchKernelGGL((batch_max_kernel<a_, b_>), dim3(b.ntiles), dim3(256), 0, s, b); break;
        DENSE_TILE_LIST(X)
#undef X
    default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

}  // namespace dense

using namespace dense;

struct dense_batch {
    std::vector<dense_cfg> cfg;
    int n = 0, scheme = -1, stages = 3;
    hipStream_t stream = nullptr;
    double* pool = nullptr;  // all instances' U^n, U1, U2
    std::vector<long long> off;
    std::vector<BatchInst> inst;
    BatchInst* d_inst = nullptr;
    BatchInstState* d_ist = nullptr;
    BatchState* d_bst = nullptr;
    int* d_tile_inst = nullptr;
    double* d_hist = nullptr;
    int hist_cap = 0;
    int tile = 0, order = 0, grid_pct = 100;
    int ntiles = 0;
    bool need_max = true;
    hipGraphExec_t graph = nullptr;      // BATCH_CHUNK steps
    int graph_steps = 0;
    hipGraphExec_t graph_rem[16] = {};   // [r]: r < BATCH_CHUNK steps, the remainder of a run
    int graph_rem_steps[16] = {};
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    double last_ms = 0.0;
    std::vector<BatchInstState> ist_host;
    char err[512] = {0};
};

static thread_local char g_berr[512];

static int bfail(dense_batch* b, int code, const char* fmt, ...)
{
    char* buf = b ? b->err : g_berr;
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, 512, fmt, ap);
    va_end(ap);
    return code;
}

#define BCHK(b, call)                                                                              \
    do {                                                                                           \
        hipError_t e_ = (call);                                                                    \
        if (e_ != hipSuccess) return bfail(b, DENSE_ERR_HIP, "%s: %s", #call, hipGetErrorString(e_)); \
    } while (0)

static void batch_release_graph(dense_batch* b)
{
    if (b->graph) (void)hipGraphExecDestroy(b->graph);
    b->graph = nullptr;
    b->graph_steps = 0;
    for (int r = 0; r < 16; r++) {
        if (b->graph_rem[r]) (void)hipGraphExecDestroy(b->graph_rem[r]);
        b->graph_rem[r] = nullptr;
        b->graph_rem_steps[r] = 0;
    }
}

// Tile list for the current (tile shape, instance order): instances in the
// chosen order, each owning a contiguous run of tiles.
static int batch_build_tiles(dense_batch* b)
{
    std::vector<int> perm(b->n);
    std::iota(perm.begin(), perm.end(), 0);
    if (b->order == 1)  // largest blocks first
        std::stable_sort(perm.begin(), perm.end(), [&](int x, int y) {
            return (long long)b->cfg[x].ni * b->cfg[x].nj > (long long)b->cfg[y].ni * b->cfg[y].nj;
        });
    std::vector<int> tile_inst;
    for (int k = 0; k < b->n; k++) {
        const int i = perm[k];
        int tx = 0;
        const int nt = tile_count(b->tile, b->cfg[i].ni, b->cfg[i].nj, &tx);
        b->inst[i].tiles_x = tx;
        b->inst[i].ntiles = nt;
        b->inst[i].tile_base = (int)tile_inst.size();
        tile_inst.insert(tile_inst.end(), nt, i);
    }
    b->ntiles = (int)tile_inst.size();
    if (b->d_tile_inst) (void)hipFree(b->d_tile_inst);
    b->d_tile_inst = nullptr;
    BCHK(b, hipMalloc(&b->d_tile_inst, sizeof(int) * tile_inst.size()));
    BCHK(b, hipMemcpy(b->d_tile_inst, tile_inst.data(), sizeof(int) * tile_inst.size(), hipMemcpyHostToDevice));
    BCHK(b, hipMemcpy(b->d_inst, b->inst.data(), sizeof(BatchInst) * b->n, hipMemcpyHostToDevice));
    batch_release_graph(b);
    b->need_max = true;
    return DENSE_OK;
}

static BatchArgs batch_args(const dense_batch* b)
{
    BatchArgs a{};
    a.inst = b->d_inst;
    a.ist = b->d_ist;
    a.bst = b->d_bst;
    a.tile_inst = b->d_tile_inst;
    a.dt_hist = b->d_hist;
    a.hist_cap = b->hist_cap;
    a.ninst = b->n;
    a.ntiles = b->ntiles;
    a.nstages = b->stages;
    a.sqrt3 = std::sqrt(3.0);
    return a;
}

// Fresh reduction ring and the wave speeds of the uploaded states.
static int batch_prime(dense_batch* b)
{
    BCHK(b, hipStreamSynchronize(b->stream));
    std::vector<BatchInstState> h(b->n);
    BCHK(b, hipMemcpy(h.data(), b->d_ist, sizeof(BatchInstState) * b->n, hipMemcpyDeviceToHost));
    for (auto& s : h) {
        std::memset(s.alpha, 0, sizeof s.alpha);
        s.error = 0;
    }
    BCHK(b, hipMemcpy(b->d_ist, h.data(), sizeof(BatchInstState) * b->n, hipMemcpyHostToDevice));
    BatchState bs{};
    BCHK(b, hipMemcpy(&bs, b->d_bst, sizeof bs, hipMemcpyDeviceToHost));
    bs.seq = 0;
    bs.blocks_done.v = 0;
    for (auto& q : bs.qnext) q.v = 0;
    bs.active = 0;
    BCHK(b, hipMemcpy(b->d_bst, &bs, sizeof bs, hipMemcpyHostToDevice));
    BatchArgs a = batch_args(b);
    BCHK(b, launch_batch_max_t(b->tile, a, b->stream));
    hipLaunchKernelGGL(batch_advance_seq_kernel, dim3(1), dim3(256), 0, b->stream, b->d_bst, b->d_ist, b->n);
    BCHK(b, hipGetLastError());
    BCHK(b, hipStreamSynchronize(b->stream));
    b->need_max = false;
    return DENSE_OK;
}

static int batch_enqueue_step(dense_batch* b)
{
    BatchArgs a = batch_args(b);
    for (int s = 1; s <= b->stages; s++) {
        a.stage = s;
        a.last_stage = (s == b->stages);
        BCHK(b, launch_batch_stage(b->scheme, b->tile, b->grid_pct, a, b->stream));
    }
    return DENSE_OK;
}

extern "C" int dense_batch_create(const dense_cfg* cfgs, int n, dense_batch_handle* out)
{
    if (!cfgs || n < 1 || !out) return bfail(nullptr, DENSE_ERR_ARG, "bad arguments");
    *out = nullptr;
    dense_batch* b = new (std::nothrow) dense_batch;
    if (!b) return bfail(nullptr, DENSE_ERR_NOMEM, "host allocation failed");
    auto bail = [&](int code) {
        dense_batch_destroy(b);
        return code;
    };
    b->n = n;
    b->cfg.assign(cfgs, cfgs + n);
    for (int i = 0; i < n; i++) {
        dense_cfg& c = b->cfg[i];
        if (c.benchmark != 0) {
            dense_cfg d = c;
            int rc = dense_problem_defaults(c.benchmark, c.ni, c.nj, &d);
            if (rc) return bail(bfail(nullptr, rc, "instance %d: bad benchmark", i));
            for (int s = 0; s < 4; s++) {
                c.bc[s] = d.bc[s];
                for (int v = 0; v < 4; v++) c.bc_state[s][v] = d.bc_state[s][v];
            }
            if (!(c.dx > 0)) c.dx = d.dx;
            if (!(c.dy > 0)) c.dy = d.dy;
        }
        char why[256];
        int rc = validate_cfg(&c, why, sizeof why);
        if (rc) return bail(bfail(nullptr, rc, "instance %d: %s", i, why));
        if (c.bc[DENSE_SIDE_LEFT] == DENSE_BC_SSL_INLET)  // needs host math at every step
            return bail(bfail(nullptr, DENSE_ERR_UNSUPPORTED, "instance %d: shear-layer inflow not batched", i));
        const int sc = scheme_of(c.splitting, c.recon);
        if (i == 0) {
            b->scheme = sc;
            b->stages = c.stages;
        } else if (sc != b->scheme || c.stages != b->stages) {
            return bail(bfail(nullptr, DENSE_ERR_ARG, "instance %d: all instances must share the scheme and stages", i));
        }
    }
    if (b->stages < 2) return bail(bfail(nullptr, DENSE_ERR_UNSUPPORTED, "batched mode needs stages >= 2"));
    // the batch's own schedule (shapes 0..3, 0 = 16x15); cfg.tile 5 maps to 16x15 too
    b->tile = (cfgs[0].tile == 4 || cfgs[0].tile == 5) ? 0 : cfgs[0].tile;
    if (hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking) != hipSuccess ||
        hipEventCreate(&b->ev0) != hipSuccess || hipEventCreate(&b->ev1) != hipSuccess)
        return bail(bfail(nullptr, DENSE_ERR_HIP, "stream/event creation failed"));
    b->inst.resize(n);
    b->off.resize(n);
    long long total = 0;
    for (int i = 0; i < n; i++) {
        const dense_cfg& c = b->cfg[i];
        BatchInst& I = b->inst[i];
        std::memset(&I, 0, sizeof I);
        I.ni = c.ni;
        I.nj = c.nj;
        I.ld = (c.ni + 31) / 32 * 32;
        I.plane = (long long)I.ld * c.nj;
        I.dx = c.dx;
        I.dy = c.dy;
        I.gamma = c.gamma;
        I.cfl = c.cfl;
        I.delta = c.delta;
        I.dt_const = c.dt;
        I.if_const_dt = c.if_const_dt;
        I.if_bodyforce = c.if_bodyforce;
        bc_fill(I.bc, c.bc, c.bc_state, c.gamma);
        b->off[i] = total;
        total += 3 * 4 * I.plane;
    }
    if (hipMalloc(&b->pool, sizeof(double) * total) != hipSuccess)
        return bail(bfail(nullptr, DENSE_ERR_NOMEM, "state pool allocation failed (%lld doubles)", total));
    if (hipMemset(b->pool, 0, sizeof(double) * total) != hipSuccess)
        return bail(bfail(nullptr, DENSE_ERR_HIP, "memset failed"));
    for (int i = 0; i < n; i++)
        for (int k = 0; k < 3; k++) b->inst[i].U[k] = b->pool + b->off[i] + k * 4 * b->inst[i].plane;
    b->hist_cap = 256;
    if (hipMalloc(&b->d_inst, sizeof(BatchInst) * n) != hipSuccess ||
        hipMalloc(&b->d_ist, sizeof(BatchInstState) * n) != hipSuccess ||
        hipMalloc(&b->d_bst, sizeof(BatchState)) != hipSuccess ||
        hipMalloc(&b->d_hist, sizeof(double) * (size_t)b->hist_cap * n) != hipSuccess)
        return bail(bfail(nullptr, DENSE_ERR_NOMEM, "batch state allocation failed"));
    std::vector<BatchInstState> z(n);
    std::memset(z.data(), 0, sizeof(BatchInstState) * n);
    BatchState bz{};
    if (hipMemcpy(b->d_ist, z.data(), sizeof(BatchInstState) * n, hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(b->d_bst, &bz, sizeof bz, hipMemcpyHostToDevice) != hipSuccess)
        return bail(bfail(nullptr, DENSE_ERR_HIP, "state init failed"));
    int rc = batch_build_tiles(b);
    if (rc) return bail(bfail(nullptr, rc, "%s", b->err));
    *out = b;
    return DENSE_OK;
}

extern "C" int dense_batch_destroy(dense_batch_handle b)
{
    if (!b) return DENSE_OK;
    if (b->stream) (void)hipStreamSynchronize(b->stream);
    batch_release_graph(b);
    if (b->pool) (void)hipFree(b->pool);
    if (b->d_inst) (void)hipFree(b->d_inst);
    if (b->d_ist) (void)hipFree(b->d_ist);
    if (b->d_bst) (void)hipFree(b->d_bst);
    if (b->d_tile_inst) (void)hipFree(b->d_tile_inst);
    if (b->d_hist) (void)hipFree(b->d_hist);
    if (b->ev0) (void)hipEventDestroy(b->ev0);
    if (b->ev1) (void)hipEventDestroy(b->ev1);
    if (b->stream) (void)hipStreamDestroy(b->stream);
    delete b;
    return DENSE_OK;
}

extern "C" int dense_batch_set_state(dense_batch_handle b, int i, const double* cons, double t0)
{
    if (!b || !cons || i < 0 || i >= b->n) return bfail(b, DENSE_ERR_ARG, "bad arguments");
    const BatchInst& I = b->inst[i];
    const int nip = I.ni + 2 * BF, njp = I.nj + 2 * BF;
    std::vector<double> h((size_t)4 * I.plane, 0.0);
    for (int k = 0; k < 4; k++)
        for (int j = 0; j < I.nj; j++)
            std::memcpy(&h[(size_t)k * I.plane + (size_t)j * I.ld], cons + (size_t)k * nip * njp + (size_t)(j + BF) * nip + BF,
                        sizeof(double) * I.ni);
    BCHK(b, hipStreamSynchronize(b->stream));
    BCHK(b, hipMemcpy(I.U[0], h.data(), sizeof(double) * h.size(), hipMemcpyHostToDevice));
    BatchInstState s{};
    s.t = t0;
    BCHK(b, hipMemcpy(&b->d_ist[i], &s, sizeof s, hipMemcpyHostToDevice));
    b->need_max = true;
    return DENSE_OK;
}

extern "C" int dense_batch_get_state(dense_batch_handle b, int i, double* cons)
{
    if (!b || !cons || i < 0 || i >= b->n) return bfail(b, DENSE_ERR_ARG, "bad arguments");
    const BatchInst& I = b->inst[i];
    const int nip = I.ni + 2 * BF, njp = I.nj + 2 * BF;
    std::vector<double> h((size_t)4 * I.plane);
    BCHK(b, hipStreamSynchronize(b->stream));
    BCHK(b, hipMemcpy(h.data(), I.U[0], sizeof(double) * h.size(), hipMemcpyDeviceToHost));
    std::memset(cons, 0, sizeof(double) * 4 * nip * njp);
    for (int k = 0; k < 4; k++)
        for (int j = 0; j < I.nj; j++)
            std::memcpy(cons + (size_t)k * nip * njp + (size_t)(j + BF) * nip + BF, &h[(size_t)k * I.plane + (size_t)j * I.ld],
                        sizeof(double) * I.ni);
    return DENSE_OK;
}

// ---------------------------------------------------------------- initial states
// Each expression is evaluated in the order written (no contraction), so the
// only roundings that can differ from a host evaluation of the same formulas
// are those of exp and pow.

// Remainder with the sign of the divisor (floor-modulo), exact like fmod.
__device__ __forceinline__ double floor_mod(double a, double b)
{
    double m = fmod(a, b);
    if (m != 0.0) {
        if ((b < 0.0) != (m < 0.0)) m += b;
    } else {
        m = copysign(0.0, b);
    }
    return m;
}

__device__ __forceinline__ void cons_of(double rho, double u, double v, double p, double gamma, double* q)
{
#pragma clang fp contract(off)  // initial states as the host evaluates them
    q[0] = rho;
    q[1] = rho * u;
    q[2] = rho * v;
    q[3] = p / (gamma - 1.0) + 0.5 * rho * (u * u + v * v);
}

__device__ __forceinline__ void ic_vortex(const dense_ic& P, double gamma, double x, double y, double* q)
{
#pragma clang fp contract(off)  // initial states as the host evaluates them
    constexpr double pi = 3.141592653589793;
    const double L = P.length;
    const double dxp = floor_mod(x - P.xc + 0.5 * L, L) - 0.5 * L;
    const double dyp = floor_mod(y - P.yc + 0.5 * L, L) - 0.5 * L;
    const double r2 = dxp * dxp + dyp * dyp;
    const double f = P.eps / (2.0 * pi) * exp(0.5 * (1.0 - r2));
    const double u = P.uinf - f * dyp;
    const double v = P.vinf + f * dxp;
    const double T = 1.0 - (gamma - 1.0) * P.eps * P.eps * exp(1.0 - r2) / (8.0 * gamma * (pi * pi));
    const double rho = pow(T, 1.0 / (gamma - 1.0));
    const double p = pow(rho, gamma);
    cons_of(rho, u, v, p, gamma, q);
}

__device__ __forceinline__ void ic_shock_vortex(const dense_ic& P, double gamma, double x, double y, double* q)
{
#pragma clang fp contract(off)  // initial states as the host evaluates them
    const double sg = sqrt(gamma);
    const double m2 = P.mach * P.mach;
    if (x < P.xs) {
        const double ddx = x - P.xc, ddy = y - P.yc;
        double r = sqrt(ddx * ddx + ddy * ddy);
        r = r > 1e-12 ? r : 1e-12;
        const double tau = r / P.rc;
        const double e2 = P.eps * P.eps;
        const double rho = pow(1.0 - (gamma - 1.0) * e2 * exp(2.0 * P.alpha * (1.0 - tau * tau)) / (4.0 * P.alpha * gamma),
                               1.0 / (gamma - 1.0));
        const double sw = P.eps * tau * exp(P.alpha * (1.0 - tau * tau));
        const double u = P.mach * sg + sw * ddy / r;
        const double v = -sw * ddx / r;
        cons_of(rho, u, v, pow(rho, gamma), gamma, q);
    } else {
        // Rankine-Hugoniot state behind the shock (shock_relation.py:7-13)
        const double rho2 = (gamma + 1.0) * m2 / (2.0 + (gamma - 1.0) * m2);
        const double p2 = 1.0 + 2.0 * gamma / (gamma + 1.0) * (m2 - 1.0);
        const double u2 = (2.0 + (gamma - 1.0) * m2) / ((gamma + 1.0) * m2) * (P.mach * sg);
        cons_of(rho2, u2, 0.0, p2, gamma, q);
    }
}

// The reference's own shock/vortex initial state, init_shockvortex_general.f90:34-56,
// operation for operation: its single-precision literals widened as Fortran
// widens them (centre 0.25, 0.5; eps 0.3, r_c 0.05, alpha 0.204; Mach 1.1; the
// cell centre (i - 0.5) formed in single precision, :21-22) and the printed
// post-shock constants (:47-50) instead of the Rankine-Hugoniot relations.
// (i, j) are 0-based interior indices.
__device__ __forceinline__ void ic_shock_vortex_ref(double gamma, int i, int j, double dx, double dy, double* q)
{
#pragma clang fp contract(off)  // initial states as the host evaluates them
    const double x_c = (double)0.25f, y_c = (double)0.5f;
    const double epsl = (double)0.3f, r_c = (double)0.05f, alpha = (double)0.204f;
    const double x = (double)((float)(i + 1) - 0.5f) * dx;
    const double y = (double)((float)(j + 1) - 0.5f) * dy;
    const double r = sqrt((x - x_c) * (x - x_c) + (y - y_c) * (y - y_c));
    const double tau = r / r_c;
    double rho, u, v, p;
    if (x < (double)0.5f) {
        rho = pow(1.0 - (gamma - 1.0) * (epsl * epsl) * exp(2.0 * alpha * (1.0 - tau * tau)) / (4.0 * alpha * gamma),
                  1.0 / (gamma - 1.0));
        u = (double)1.1f * sqrt(gamma) + epsl * tau * exp(alpha * (1.0 - tau * tau)) * (y - y_c) / r;
        v = -(epsl * tau * exp(alpha * (1.0 - tau * tau)) * (x - x_c) / r);
        p = pow(rho, gamma);
    } else {
        rho = 1.169082;
        u = 1.113299;
        v = 0.0;
        p = 1.245000;
    }
    cons_of(rho, u, v, p, gamma, q);
}

// blockIdx.y = instance; the x blocks stride over its whole plane (row pitch
// included), so padding columns are written as zeros like dense_batch_set_state.
__global__ void __launch_bounds__(256) batch_ic_kernel(const BatchInst* __restrict__ inst,
                                                       const dense_ic* __restrict__ ics)
{
    const BatchInst& I = inst[blockIdx.y];
    const dense_ic P = ics[blockIdx.y];
    const long long n = I.plane;
    double* U = I.U[0];
    for (long long c = (long long)blockIdx.x * 256 + threadIdx.x; c < n; c += (long long)gridDim.x * 256) {
        const int j = (int)(c / I.ld), i = (int)(c - (long long)j * I.ld);
        double q[4] = {0.0, 0.0, 0.0, 0.0};
        if (i < I.ni) {
            const double x = ((double)i + 0.5) * I.dx, y = ((double)j + 0.5) * I.dy;
            if (P.kind == DENSE_IC_VORTEX)
                ic_vortex(P, I.gamma, x, y, q);
            else if (P.kind == DENSE_IC_SHOCK_VORTEX)
                ic_shock_vortex(P, I.gamma, x, y, q);
            else
                ic_shock_vortex_ref(I.gamma, i, j, I.dx, I.dy, q);
        }
#pragma unroll
        for (int k = 0; k < 4; k++) U[k * n + c] = q[k];
    }
}

extern "C" int dense_batch_init_ic(dense_batch_handle b, const dense_ic* ics)
{
    if (!b || !ics) return bfail(b, DENSE_ERR_ARG, "bad arguments");
    long long maxplane = 0;
    for (int i = 0; i < b->n; i++) {
        const dense_ic& P = ics[i];
        const double v[] = {P.eps, P.xc, P.yc, P.uinf, P.vinf, P.length, P.mach, P.xs, P.rc, P.alpha};
        for (double x : v)
            if (!std::isfinite(x)) return bfail(b, DENSE_ERR_ARG, "instance %d: non-finite parameter", i);
        if (P.kind == DENSE_IC_VORTEX) {
            if (!(P.length > 0.0)) return bfail(b, DENSE_ERR_ARG, "instance %d: vortex needs length > 0", i);
        } else if (P.kind == DENSE_IC_SHOCK_VORTEX) {
            if (!(P.mach >= 1.0) || !(P.rc > 0.0))
                return bfail(b, DENSE_ERR_ARG, "instance %d: shock/vortex needs mach >= 1 and rc > 0", i);
        } else if (P.kind != DENSE_IC_SHOCK_VORTEX_REF) {
            return bfail(b, DENSE_ERR_ARG, "instance %d: unknown initial-state kind %d", i, P.kind);
        }
        maxplane = std::max(maxplane, b->inst[i].plane);
    }
    if (b->n > 65535) return bfail(b, DENSE_ERR_ARG, "at most 65535 instances");
    dense_ic* d_ic = nullptr;
    BCHK(b, hipStreamSynchronize(b->stream));
    if (hipMalloc(&d_ic, sizeof(dense_ic) * b->n) != hipSuccess)
        return bfail(b, DENSE_ERR_NOMEM, "descriptor allocation failed");
    hipError_t e = hipMemcpy(d_ic, ics, sizeof(dense_ic) * b->n, hipMemcpyHostToDevice);
    if (e == hipSuccess) {
        const long long blocks = std::min<long long>((maxplane + 255) / 256, 64);
        hipLaunchKernelGGL(batch_ic_kernel, dim3((unsigned)blocks, (unsigned)b->n), dim3(256), 0, b->stream, b->d_inst,
                           d_ic);
        e = hipGetLastError();
    }
    if (e == hipSuccess) e = hipMemsetAsync(b->d_ist, 0, sizeof(BatchInstState) * b->n, b->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(b->stream);
    (void)hipFree(d_ic);
    if (e != hipSuccess) return bfail(b, DENSE_ERR_HIP, "initial states: %s", hipGetErrorString(e));
    b->need_max = true;
    return DENSE_OK;
}

constexpr int BATCH_CHUNK = 16;
static_assert(BATCH_CHUNK <= 16, "one cached remainder graph per count below the chunk");

// Capture nsteps steps into *exec (replacing what it held).
static int batch_build_graph(dense_batch* b, int nsteps, hipGraphExec_t* exec, int* exec_steps)
{
    if (*exec) (void)hipGraphExecDestroy(*exec);
    *exec = nullptr;
    *exec_steps = 0;
    hipGraph_t g = nullptr;
    BCHK(b, hipStreamBeginCapture(b->stream, hipStreamCaptureModeThreadLocal));
    for (int i = 0; i < nsteps; i++) {
        int rc = batch_enqueue_step(b);
        if (rc) {
            (void)hipStreamEndCapture(b->stream, &g);
            if (g) (void)hipGraphDestroy(g);
            return rc;
        }
    }
    BCHK(b, hipStreamEndCapture(b->stream, &g));
    hipError_t e = hipGraphInstantiate(exec, g, nullptr, nullptr, 0);
    (void)hipGraphDestroy(g);
    if (e != hipSuccess) {
        *exec = nullptr;
        return bfail(b, DENSE_ERR_HIP, "hipGraphInstantiate: %s", hipGetErrorString(e));
    }
    *exec_steps = nsteps;
    return DENSE_OK;
}

extern "C" int dense_batch_run(dense_batch_handle b, int nsteps, double* dt_hist, double* t_out)
{
    if (!b || nsteps < 0) return bfail(b, DENSE_ERR_ARG, "bad arguments");
    if (nsteps == 0) return DENSE_OK;
    if (b->need_max) {
        int rc = batch_prime(b);
        if (rc) return rc;
    }
    if (dt_hist && nsteps > b->hist_cap) {
        // the device history grows to the longest run a caller asks a history
        // for; the captured graph carries its address and is captured again
        BCHK(b, hipStreamSynchronize(b->stream));
        double* p = nullptr;
        if (hipMalloc(&p, sizeof(double) * (size_t)nsteps * b->n) != hipSuccess)
            return bfail(b, DENSE_ERR_NOMEM, "dt history of %d steps: allocation failed", nsteps);
        (void)hipFree(b->d_hist);
        b->d_hist = p;
        b->hist_cap = nsteps;
        batch_release_graph(b);
    }
    // whole chunks, then one graph of the remainder: no replayed step is idle
    const int full = nsteps / BATCH_CHUNK, rem = nsteps % BATCH_CHUNK;
    if (full > 0 && b->graph_steps != BATCH_CHUNK) {
        int rc = batch_build_graph(b, BATCH_CHUNK, &b->graph, &b->graph_steps);
        if (rc) return rc;
    }
    if (rem > 0 && b->graph_rem_steps[rem] != rem) {
        int rc = batch_build_graph(b, rem, &b->graph_rem[rem], &b->graph_rem_steps[rem]);
        if (rc) return rc;
    }
    // clear the per-instance step counters used to index the dt history
    std::vector<BatchInstState> h(b->n);
    BCHK(b, hipMemcpy(h.data(), b->d_ist, sizeof(BatchInstState) * b->n, hipMemcpyDeviceToHost));
    for (auto& s : h) s.step = 0;
    BCHK(b, hipMemcpy(b->d_ist, h.data(), sizeof(BatchInstState) * b->n, hipMemcpyHostToDevice));
    BatchState bs{};
    BCHK(b, hipMemcpy(&bs, b->d_bst, sizeof bs, hipMemcpyDeviceToHost));
    bs.step = 0;
    BCHK(b, hipMemcpy(b->d_bst, &bs, sizeof bs, hipMemcpyHostToDevice));
    hipLaunchKernelGGL(batch_set_run_kernel, dim3(1), dim3(1), 0, b->stream, b->d_bst, nsteps);
    BCHK(b, hipGetLastError());
    BCHK(b, hipEventRecord(b->ev0, b->stream));
    for (int i = 0; i < full; i++) BCHK(b, hipGraphLaunch(b->graph, b->stream));
    if (rem > 0) BCHK(b, hipGraphLaunch(b->graph_rem[rem], b->stream));
    BCHK(b, hipEventRecord(b->ev1, b->stream));
    BCHK(b, hipStreamSynchronize(b->stream));
    float ms = 0.f;
    BCHK(b, hipEventElapsedTime(&ms, b->ev0, b->ev1));
    b->last_ms = ms;
    BCHK(b, hipMemcpy(h.data(), b->d_ist, sizeof(BatchInstState) * b->n, hipMemcpyDeviceToHost));
    int err = -1;
    for (int i = 0; i < b->n; i++) {
        if (t_out) t_out[i] = h[i].t;
        if (h[i].error && err < 0) err = i;
    }
    if (dt_hist) BCHK(b, hipMemcpy(dt_hist, b->d_hist, sizeof(double) * (size_t)nsteps * b->n, hipMemcpyDeviceToHost));
    if (err >= 0) return bfail(b, DENSE_ERR_NONFINITE, "instance %d: non-finite wave speed", err);
    return DENSE_OK;
}

extern "C" int dense_batch_set_schedule(dense_batch_handle b, int tile, int order, int grid_pct)
{
    if (!b || tile < 0 || tile > 3 || order < 0 || order > 1 || grid_pct < 1 || grid_pct > 100)
        return bfail(b, DENSE_ERR_ARG, "bad schedule");
    BCHK(b, hipStreamSynchronize(b->stream));
    b->tile = tile;
    b->order = order;
    b->grid_pct = grid_pct;
    return batch_build_tiles(b);
}

// Time each (tile shape, instance order, grid fraction) on `trial_steps` steps
// of the current states, restore the states after every trial, keep the
// fastest.  All candidates give bit-identical results.
extern "C" int dense_batch_autotune(dense_batch_handle b, int trial_steps, int* tile, int* order, int* grid_pct,
                                    double* best_ms)
{
    if (!b || trial_steps < 1) return bfail(b, DENSE_ERR_ARG, "bad arguments");
    BCHK(b, hipStreamSynchronize(b->stream));
    long long total = 0;
    for (int i = 0; i < b->n; i++) total += 4 * b->inst[i].plane;
    // save U^n of every instance and the clocks (on failure: free what was
    // taken, states and schedule untouched)
    std::vector<double*> save(b->n, nullptr);
    std::vector<BatchInstState> ist0(b->n);
    const int t_in = b->tile, o_in = b->order, g_in = b->grid_pct;
    int rc = DENSE_OK;
    for (int i = 0; i < b->n && !rc; i++) {
        if (hipMalloc(&save[i], sizeof(double) * 4 * b->inst[i].plane) != hipSuccess)
            rc = bfail(b, DENSE_ERR_NOMEM, "autotune: state snapshot allocation failed");
        else if (hipMemcpy(save[i], b->inst[i].U[0], sizeof(double) * 4 * b->inst[i].plane,
                           hipMemcpyDeviceToDevice) != hipSuccess)
            rc = bfail(b, DENSE_ERR_HIP, "autotune: state snapshot failed");
    }
    if (!rc && hipMemcpy(ist0.data(), b->d_ist, sizeof(BatchInstState) * b->n, hipMemcpyDeviceToHost) != hipSuccess)
        rc = bfail(b, DENSE_ERR_HIP, "autotune: clock snapshot failed");
    if (rc) {
        for (double* p : save)
            if (p) (void)hipFree(p);
        return rc;
    }
    auto restore = [&]() -> int {
        for (int i = 0; i < b->n; i++)
            BCHK(b, hipMemcpy(b->inst[i].U[0], save[i], sizeof(double) * 4 * b->inst[i].plane, hipMemcpyDeviceToDevice));
        BCHK(b, hipMemcpy(b->d_ist, ist0.data(), sizeof(BatchInstState) * b->n, hipMemcpyHostToDevice));
        b->need_max = true;
        return DENSE_OK;
    };
    double best = 1e300;
    int bt = b->tile, bo = b->order, bg = b->grid_pct;
    const int tiles[] = {0, 1, 2, 3};
    const int orders[] = {0, 1};
    const int grids[] = {100, 50};
    for (int t : tiles)
        for (int o : orders)
            for (int g : grids) {
                if ((rc = dense_batch_set_schedule(b, t, o, g))) goto done;
                if ((rc = dense_batch_run(b, trial_steps, nullptr, nullptr))) goto done;  // includes graph build
                if ((rc = restore())) goto done;
                if ((rc = dense_batch_run(b, trial_steps, nullptr, nullptr))) goto done;  // timed
                if (b->last_ms < best) {
                    best = b->last_ms;
                    bt = t, bo = o, bg = g;
                }
                if ((rc = restore())) goto done;
            }
    rc = dense_batch_set_schedule(b, bt, bo, bg);
done:
    if (rc) {
        // a failed trial: the states and clocks as they were, the caller's schedule
        const int keep = rc;
        (void)restore();
        (void)dense_batch_set_schedule(b, t_in, o_in, g_in);
        rc = keep;
    }
    for (double* p : save)
        if (p) (void)hipFree(p);
    if (rc) return rc;
    if (tile) *tile = bt;
    if (order) *order = bo;
    if (grid_pct) *grid_pct = bg;
    if (best_ms) *best_ms = best;
    return DENSE_OK;
}

extern "C" int dense_batch_last_run_gpu_ms(dense_batch_handle b, double* ms)
{
    if (!b || !ms) return bfail(b, DENSE_ERR_ARG, "null argument");
    *ms = b->last_ms;
    return DENSE_OK;
}

extern "C" const char* dense_batch_last_error(dense_batch_handle b) { return b ? b->err : g_berr; }
