// Tile machinery shared by the single-block and the batched RK-stage kernels.
//
// One workgroup owns a TX x TY tile of interior cells.  Per tile and stage it
//   1. stages the tile plus a 3-cell halo of conserved variables in LDS with
//      coalesced loads along i, deriving primitives on the fly
//      (get_primary.f90:46-67) and synthesising ghost cells from the boundary
//      rules (boundary_*_general.f90) so no ghost layer lives in HBM;
//   2. reconstructs all x faces, then all y faces of the tile (one face per
//      thread), each face producing h = h+ + h- (flux_splitting_*.f90,
//      scheme_*.f90, field_op.f90:1-21);
//   3. applies the SSP-RK3 stage update (time_advance_general.f90) and
//      folds max|u|+c, max|v|+c of the new state into the thread's running max
//      for the next stage's global Lax-Friedrichs speed and CFL step.
// Nothing here is an MFMA contraction; the kernels are FP64-VALU bound.
#pragma once
#include "dense_faces.h"

namespace dense {

template <int TX_, int TY_, int NT_>
struct Tile {
    static constexpr int TX = TX_, TY = TY_, NT = NT_;
    static constexpr int W = TX + 6, H = TY + 6, NC = W * H;
    static constexpr int NXF = (TX + 1) * TY, NYF = TX * (TY + 1);
    static_assert(NXF <= NT && NYF <= NT, "one face per thread per direction");
    static_assert(4 * (NXF + NYF) <= REC_N * NC, "h arrays alias the cell record");
    // DIRSPLIT: threads [0, NT/2) take the x faces and [NT/2, NT) the y faces
    // (one face per thread instead of one per direction).
    static constexpr bool DIRSPLIT = NXF <= NT / 2 && NYF <= NT / 2;
};

// Tile shapes (index = dense_cfg.tile).  Chosen so that each direction has
// at most one face per thread and a multiple of 64 threads.
#define DENSE_TILE_LIST(X) \
    X(0, 16, 15, 256)      \
    X(1, 32, 7, 256)       \
    X(2, 8, 7, 64)         \
    X(3, 32, 15, 512)      \
    X(4, 16, 15, 512)

constexpr int NTILES_SHAPES = 5;

__host__ __device__ inline void tile_shape(int tile, int& tx, int& ty, int& nt)
{
    switch (tile) {
#define X(i, a, b, c) \
    case i: tx = a; ty = b; nt = c; return;
        DENSE_TILE_LIST(X)
#undef X
    default: tx = 16; ty = 15; nt = 256; return;
    }
}

// Block -> tile map that hands each XCD a contiguous run of tiles (workgroups
// are dealt round-robin over the 8 XCDs), so halo re-reads hit the same L2.
__device__ __forceinline__ int xcd_tile(int b, int n)
{
    const int q = n >> 3, r = n & 7;
    const int x = b & 7, k = b >> 3;
    return x * q + (x < r ? x : r) + k;
}

__device__ __forceinline__ unsigned long long dbits(double x) { return (unsigned long long)__double_as_longlong(x); }

__device__ __forceinline__ unsigned long long wave_max_u64(unsigned long long v)
{
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) {
        const unsigned long long o = __shfl_xor(v, off, 64);
        v = o > v ? o : v;
    }
    return v;
}

// s_waitcnt vmcnt(0): every vector-memory operation of this wave, stores and
// atomics without return included, has been acknowledged by the memory
// system (expcnt and lgkmcnt left unconstrained: gfx9 encoding).
__device__ __forceinline__ void wait_vm_writes() { __builtin_amdgcn_s_waitcnt(0x0F70); }

// Block-wide max of (ei, ej) bit patterns into the out slot.  Launches of up
// to 512 workgroups apply both atomics at once: reading the slot first is a
// memory round trip that each workgroup would wait for at its end (measured
// 7 % of a 40,000-cell launch, where all workgroups end together).  Larger
// launches read both slot values at once and skip atomics that cannot raise
// them: unconditional atomics from 768 workgroups queue on the slot's line
// (3 % of a 921,600-cell LLF launch).
template <int NT>
__device__ __forceinline__ void block_max_to_slot(unsigned long long ei, unsigned long long ej, unsigned long long* slot)
{
    __shared__ unsigned long long red[2][NT / 64];
    ei = wave_max_u64(ei);
    ej = wave_max_u64(ej);
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    if (lane == 0) {
        red[0][w] = ei;
        red[1][w] = ej;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        unsigned long long mi = red[0][0], mj = red[1][0];
#pragma unroll
        for (int k = 1; k < NT / 64; k++) {
            mi = red[0][k] > mi ? red[0][k] : mi;
            mj = red[1][k] > mj ? red[1][k] : mj;
        }
        if (gridDim.x <= 512) {
            atomicMax(&slot[0], mi);
            atomicMax(&slot[1], mj);
        } else {
            const unsigned long long ci = __hip_atomic_load(&slot[0], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            const unsigned long long cj = __hip_atomic_load(&slot[1], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            if (mi > ci) atomicMax(&slot[0], mi);
            if (mj > cj) atomicMax(&slot[1], mj);
        }
    }
}

// The last workgroup of a launch advances the kernel sequence number, clears
// the slot the next launch reduces into, checks the new wave speeds and, at
// the end of a time step, advances the device clock and the save clock
// (general_main.f90:35-37).
// Only atomics cross workgroups here.  Producer side: each block waits until
// its max atomics have been performed (wait_vm_writes: agent-scope atomics
// complete in issue order once vmcnt reaches zero) before its relaxed counter
// increment; this avoids an agent-scope release fence, which would write the
// L2 back on every block of a multi-XCD part.  Consumer side: the block that
// takes the last count issues an agent-scope acquire fence (once per launch)
// before it reads the maxima.  Stores to the state arrays and the DevState
// fields reach the next launch through the kernel boundary.
__device__ void launch_epilogue(DevState* st, int seq, bool end_of_step, double t, double dt, double* dt_hist)
{
    __syncthreads();
    if (threadIdx.x != 0) return;
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
    wait_vm_writes();
    const unsigned int prev =
        __hip_atomic_fetch_add(&st->blocks_done.v, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    if (prev != gridDim.x - 1) return;
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
    __hip_atomic_store(&st->blocks_done.v, 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
#pragma unroll
    for (int x = 0; x < 8; x++) __hip_atomic_store(&st->qnext[x].v, 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    const int out = (seq + 1) & 3, nxt = (seq + 2) & 3;
    __hip_atomic_store(&st->alpha[nxt][0], 0ull, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    __hip_atomic_store(&st->alpha[nxt][1], 0ull, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    const unsigned long long ai = __hip_atomic_load(&st->alpha[out][0], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    const unsigned long long aj = __hip_atomic_load(&st->alpha[out][1], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    const unsigned long long inf_bits = 0x7ff0000000000000ull;
    if (ai >= inf_bits || aj >= inf_bits) st->error = 1;
    st->seq = seq + 1;
    if (end_of_step) {
        if (dt_hist && st->hist_n < st->hist_cap) dt_hist[st->hist_n] = dt;
        st->hist_n += 1;
        st->t_prev = t;
        st->t = t + dt;
        st->step += 1;
        const double el = st->elasp + dt;  // elasp_time = elasp_time + dt, general_main.f90:37
        st->elasp = el;
        st->active = (st->t < st->t_end) && (st->step < st->step_end) && !(el >= st->elasp_end) && !st->error;
    }
}

// CFL step size, time_advance_general.f90:20-24
__device__ __forceinline__ double cfl_dt(const StageArgs& a, double ai, double aj)
{
    if (a.if_const_dt) return a.dt_const;
    const double dt_xi = a.dx / ai;
    const double dt_eta = a.dy / aj;
    return a.cfl * dt_xi * dt_eta / (dt_xi + dt_eta);
}

template <int SCHEME, int DIR, int NC, int ST>
__device__ __forceinline__ void eval_face(const double* __restrict__ s, int base, const StageArgs& a, double alpha,
                                          double* h)
{
    if constexpr (SCHEME == S_CH) face_char<DIR, NC, ST>(s, base, alpha, a.gamma, h);
    else if constexpr (SCHEME == S_SC) face_ada<DIR, NC, ST>(s, base, alpha, a.gamma, a.delta, h);
    else if constexpr (SCHEME >= S_SW_UP5) face_cw_sw<DIR, SCHEME - S_SW_UP5, NC, ST>(s, base, a.gamma, h);
    else if constexpr (SCHEME >= S_GLF_UP5) face_cw_lf<DIR, false, SCHEME - S_GLF_UP5, NC, ST>(s, base, alpha, h);
    else face_cw_lf<DIR, true, SCHEME - S_LLF_UP5, NC, ST>(s, base, alpha, h);
}

// Only the LLF (per-cell wave speed |u_n| + c) and Steger-Warming faces read a
// cell's sound speed; the tile loader skips it for the others.
template <int SCHEME>
constexpr bool faces_read_sound_speed()
{
    return SCHEME <= S_LLF_Z || (SCHEME >= S_SW_UP5 && SCHEME <= S_SW_Z);
}

// Minimum waves per SIMD requested from the register allocator: 3 for every
// scheme (the characteristic ones run their flux signs one after the other to
// fit), 4 for the aliased upwind5 tiles below.
#ifndef DENSE_WAVES
#define DENSE_WAVES 3
#endif
// The Lax-Friedrichs upwind5 schemes on 512-thread tiles keep a face's h in
// registers until every face of the tile has read the cell records, then write
// the h arrays over those records (one more barrier).  A 32x15 tile then takes
// 51 KB of LDS instead of 83 KB and, at 4 waves per SIMD (128 VGPRs), two
// workgroups fit on a CU instead of one.  On 256-thread tiles the same change
// measured 1-3 % slower (the 128-VGPR cap spills, and 3 workgroups of 46 KB
// already hide the load phase); the other schemes need more than 128 VGPRs,
// so their LDS never limits the workgroups per CU.
#ifndef DENSE_ALIAS_H
#define DENSE_ALIAS_H 1
#endif
template <int SCHEME, class T>
constexpr bool alias_h()
{
    return DENSE_ALIAS_H && T::NT >= 512 && (SCHEME == S_LLF_UP5 || SCHEME == S_GLF_UP5);
}

template <int SCHEME, class T>
constexpr int min_waves()
{
    return alias_h<SCHEME, T>() ? 4 : DENSE_WAVES;
}

// Tile loader of threads that hold several cells: the component-wise WENO,
// Steger-Warming and AdaWENO kernels load all of a thread's cells at once (one
// memory round trip per tile), upwind5 and the characteristic kernel one cell
// after the other, which measured faster for them.  Double Mach 1920x480,
// 16x15 tiles, us per stage launch, per cell / at once: LLF + upwind5
// 45.1 / 48.9, 104 72.0 / 72.1, AdaWENO 55.2 / 51.7, GLF + WENO-Z 61.5 / 60.5.
template <int SCHEME>
constexpr bool batch_loads()
{
    return !(SCHEME == S_LLF_UP5 || SCHEME == S_GLF_UP5 || SCHEME == S_SW_UP5 || SCHEME == S_CH);
}

// Per-XCD tile queue of a persistent launch: queue x owns the contiguous
// tile range [lo, hi); workgroups x, x+8, x+16, ... (dealt to XCD x) pull from it.
__device__ __forceinline__ void queue_range(int x, int n, int& lo, int& hi)
{
    const int q = n >> 3, r = n & 7;
    lo = x * q + (x < r ? x : r);
    hi = lo + q + (x < r ? 1 : 0);
}

// Steger-Warming: a cell's split fluxes F+ and F- (sw_flux: 10 divisions and
// 3 square roots) are computed once per sweep direction by the tile loader into
// LDS bands, instead of once for each of the six faces whose stencil holds the
// cell.  Bands: the x-split of the tile's TY rows (all W columns), the y-split
// of its TX columns (all H rows), and the conserved state of the tile's own
// cells for the update; the h arrays overwrite the bands after the faces.
#ifndef DENSE_SW_BANDS
#define DENSE_SW_BANDS 1
#endif
template <int SCHEME>
constexpr bool split_in_loader()
{
    return DENSE_SW_BANDS && SCHEME >= S_SW_UP5 && SCHEME <= S_SW_Z;
}

// LDS footprint of one tile.  Cell records [REC_N][H][W], or with BANDS the
// conserved state [4][TY][TX], the x band [8][TY][W] (F+ then F-) and the y band
// [8][H][TX].  h at x faces [4][TY][TX+1], then y faces [4][TY+1][TX]: in sh,
// over the cell records (ALIAS) or over the bands (BANDS).
template <int SCHEME, class T>
struct TileLds {
    static constexpr bool BANDS = split_in_loader<SCHEME>();
    static constexpr bool ALIAS = alias_h<SCHEME, T>();
    static constexpr int XB = T::W * T::TY, YB = T::TX * T::H;  // band plane sizes
    static constexpr int NS = BANDS ? 4 * T::TX * T::TY + 8 * (XB + YB) : REC_N * T::NC;
    static_assert(!BANDS || 4 * (T::NXF + T::NYF) <= 8 * (XB + YB), "h arrays alias the bands");
    double s[NS];
    double sh[(ALIAS || BANDS) ? 1 : 4 * (T::NXF + T::NYF)];
};

// One Steger-Warming face from the bands: the 6 split fluxes of its stencil
// are band[base + m*st], m = 0..5 (cells f-3 .. f+2), as face_cw_sw forms them.
template <int RECON>
__device__ __forceinline__ void face_sw_bands(const double* __restrict__ band, int plane, int base, int st,
                                              double* h)
{
#pragma unroll
    for (int k = 0; k < 4; k++) {
        double lfp[5], lfm[5];
#pragma unroll
        for (int i = 0; i < 5; i++) {
            lfp[i] = band[k * plane + base + i * st];
            lfm[i] = band[(4 + k) * plane + base + (5 - i) * st];
        }
        h[k] = recon<RECON>(lfp) + recon<RECON>(lfm);
    }
}

// One tile of one RK stage.  `a` describes the block (pointers of this stage,
// grid, metrics, boundaries); ai/aj are the stage's global wave speeds, dt the
// step, t the time of the stage's boundary values.  after_load() runs on
// thread 0 between the load barrier and the face phase (used to fetch the
// next tile index while the faces compute).  Ends with a barrier.
template <int SCHEME, class T, class AfterLoad>
__device__ __forceinline__ void run_tile(const StageArgs& a, int tile, double ai, double aj, double dt, double t,
                                         TileLds<SCHEME, T>& L, unsigned long long& ei,
                                         unsigned long long& ej,
                                         AfterLoad after_load)
{
    constexpr int TX = T::TX, TY = T::TY, NT = T::NT, W = T::W, NC = T::NC;
    static_assert(TX * TY <= NT, "one update cell per thread");
    double* s = L.s;
    constexpr bool ALIAS = alias_h<SCHEME, T>();
    constexpr bool BANDS = split_in_loader<SCHEME>();
    constexpr bool HOLD = ALIAS || BANDS;  // h waits in registers, then overwrites the records / bands
    constexpr int XB = TileLds<SCHEME, T>::XB, YB = TileLds<SCHEME, T>::YB;
    double* sc = L.s;                    // BANDS: conserved state of the tile's cells
    double* xs = L.s + 4 * TX * TY;      // BANDS: x band
    double* ys = xs + 8 * XB;            // BANDS: y band
    double* shx = BANDS ? xs : ALIAS ? L.s : L.sh;
    double* shy = shx + 4 * T::NXF;
    const int tid = threadIdx.x;
    const int ni = a.ni, nj = a.nj;
    const int x0 = (tile % a.tiles_x) * TX, y0 = (tile / a.tiles_x) * TY;

    // ---- 1. tile + halo -> LDS (corners and cells beyond the 3rd ghost layer are never read).
    // Cell idx = tid + r*NT of the W x H record, r < NR.
    constexpr int NR = (NC + NT - 1) / NT;
    const auto needed = [&](int idx, int& gx, int& gy) {
        const int ly = idx / W, lx = idx - ly * W;
        gx = x0 + lx - 3;
        gy = y0 + ly - 3;
        const bool inx = (unsigned)gx < (unsigned)ni, iny = (unsigned)gy < (unsigned)nj;
        if (idx >= NC || (!inx && !iny) || gx >= ni + 3 || gy >= nj + 3) return false;
        if constexpr (BANDS) {
            if ((unsigned)(ly - 3) >= (unsigned)TY && (unsigned)(lx - 3) >= (unsigned)TX) return false;
        }
        return true;
    };
    const auto inside = [&](int gx, int gy) { return (unsigned)gx < (unsigned)ni && (unsigned)gy < (unsigned)nj; };
    // The record of one cell from its loaded conserved state (interior) or the
    // boundary rules (ghost).
    const auto put_cell = [&](int idx, int gx, int gy, double* c) {
        const int ly = idx / W, lx = idx - ly * W;
        double pr[4];
        if (inside(gx, gy)) prim_of<faces_read_sound_speed<SCHEME>()>(a.gamma, c, pr);
        else ghost_cell(a, a.us, gx, gy, t, c, pr);
        if constexpr (BANDS) {
            const bool bx = (unsigned)(ly - 3) < (unsigned)TY, by = (unsigned)(lx - 3) < (unsigned)TX;
            if (bx && by) {
#pragma unroll
                for (int k = 0; k < 4; k++) sc[k * TX * TY + (ly - 3) * TX + lx - 3] = c[k];
            }
            if (bx) {
                double fp[4], fm[4];
                sw_flux<0>(a.gamma, c[0], pr[0], pr[1], pr[3], fp, fm);
#pragma unroll
                for (int k = 0; k < 4; k++) {
                    xs[k * XB + (ly - 3) * W + lx] = fp[k];
                    xs[(4 + k) * XB + (ly - 3) * W + lx] = fm[k];
                }
            }
            if (by) {
                double fp[4], fm[4];
                sw_flux<1>(a.gamma, c[0], pr[0], pr[1], pr[3], fp, fm);
#pragma unroll
                for (int k = 0; k < 4; k++) {
                    ys[k * YB + ly * TX + lx - 3] = fp[k];
                    ys[(4 + k) * YB + ly * TX + lx - 3] = fm[k];
                }
            }
        } else {
#pragma unroll
            for (int k = 0; k < 4; k++) {
                s[k * NC + idx] = c[k];
                s[(4 + k) * NC + idx] = pr[k];
            }
        }
    };
    if constexpr (NR == 1 || !batch_loads<SCHEME>()) {
        for (int idx = tid; idx < NC; idx += NT) {
            int gx, gy;
            if (!needed(idx, gx, gy)) continue;
            double c[4] = {0, 0, 0, 0};
            if (inside(gx, gy)) {
                const long long o = (long long)gy * a.ld + gx;
#pragma unroll
                for (int k = 0; k < 4; k++) c[k] = a.us[k * a.plane + o];
            }
            put_cell(idx, gx, gy, c);
        }
    } else {
        // Several cells per thread: all of them are loaded at once,
        // unconditionally, from coordinates clamped into the block (the values
        // of ghost cells and of idx >= NC are not used), then converted one
        // after the other by a single copy of the conversion code: one memory
        // round trip per tile instead of one per cell.
        double cin[NR][4];
#pragma unroll
        for (int r = 0; r < NR; r++) {
            const int idx = min(tid + r * NT, NC - 1);
            const int ly = idx / W, lx = idx - ly * W;
            const int gx = min(max(x0 + lx - 3, 0), ni - 1), gy = min(max(y0 + ly - 3, 0), nj - 1);
            const long long o = (long long)gy * a.ld + gx;
#pragma unroll
            for (int k = 0; k < 4; k++) cin[r][k] = a.us[k * a.plane + o];
        }
        // Keep the loads here: without this the compiler sinks each cell's
        // loads into the branch that converts it, one round trip per cell again.
#pragma unroll
        for (int r = 0; r < NR; r++) {
#pragma unroll
            for (int k = 0; k < 4; k++) asm volatile("" : "+v"(cin[r][k]));
        }
#pragma unroll 1
        for (int r = 0; r < NR; r++) {
            const int idx = tid + r * NT;
            int gx, gy;
            if (needed(idx, gx, gy)) {
                double c[4] = {cin[0][0], cin[0][1], cin[0][2], cin[0][3]};
                put_cell(idx, gx, gy, c);
            }
#pragma unroll
            for (int q = 0; q + 1 < NR; q++) {  // the next cell's values move to cin[0]
#pragma unroll
                for (int k = 0; k < 4; k++) cin[q][k] = cin[q + 1][k];
            }
        }
    }
    __syncthreads();
    if (tid == 0) after_load();

    // U^n of this thread's update cell, in flight during the face phase
    const int lxu = tid % TX, lyu = tid / TX;
    const int gxu = x0 + lxu, gyu = y0 + lyu;
    const bool vu = tid < TX * TY && gxu < ni && gyu < nj;
    const long long ou = (long long)gyu * a.ld + gxu;
    double q0[4] = {0, 0, 0, 0};
    if (vu && a.stage > 1) {
#pragma unroll
        for (int k = 0; k < 4; k++) q0[k] = a.un[k * a.plane + ou];
    }

    // ---- 2. faces: x faces (fx in [0,TX], ly in [0,TY)), then y faces
    const auto put_x = [&](const double* h, int o) {
#pragma unroll
        for (int k = 0; k < 4; k++) shx[k * T::NXF + o] = h[k];
    };
    const auto put_y = [&](const double* h, int o) {
#pragma unroll
        for (int k = 0; k < 4; k++) shy[k * T::NYF + o] = h[k];
    };
    double hx[8], hy[8];
    int ox = -1, oy = -1;  // this thread's x / y face in the h arrays, -1 if none
    if constexpr (T::DIRSPLIT) {
        if (tid < NT / 2) {
            const int fx = tid % (TX + 1), lyx = tid / (TX + 1);
            if (tid < T::NXF && x0 + fx <= ni && y0 + lyx < nj) {
                if constexpr (BANDS) face_sw_bands<SCHEME - S_SW_UP5>(xs, XB, lyx * W + fx, 1, hx);
                else eval_face<SCHEME, 0, NC, 1>(s, (lyx + 3) * W + fx, a, ai, hx);
                ox = lyx * (TX + 1) + fx;
                if constexpr (!HOLD) put_x(hx, ox);
            }
        } else {
            const int f = tid - NT / 2;
            const int lxy = f % TX, fy = f / TX;
            if (f < T::NYF && x0 + lxy < ni && y0 + fy <= nj) {
                if constexpr (BANDS) face_sw_bands<SCHEME - S_SW_UP5>(ys, YB, fy * TX + lxy, TX, hy);
                else eval_face<SCHEME, 1, NC, W>(s, fy * W + lxy + 3, a, aj, hy);
                oy = fy * TX + lxy;
                if constexpr (!HOLD) put_y(hy, oy);
            }
        }
    } else {
        const int fx = tid % (TX + 1), lyx = tid / (TX + 1);
        if (tid < T::NXF && x0 + fx <= ni && y0 + lyx < nj) {
            if constexpr (BANDS) face_sw_bands<SCHEME - S_SW_UP5>(xs, XB, lyx * W + fx, 1, hx);
            else eval_face<SCHEME, 0, NC, 1>(s, (lyx + 3) * W + fx, a, ai, hx);
            ox = lyx * (TX + 1) + fx;
            if constexpr (!HOLD) put_x(hx, ox);
        }
        const int lxy = tid % TX, fy = tid / TX;
        if (tid < T::NYF && x0 + lxy < ni && y0 + fy <= nj) {
            if constexpr (BANDS) face_sw_bands<SCHEME - S_SW_UP5>(ys, YB, fy * TX + lxy, TX, hy);
            else eval_face<SCHEME, 1, NC, W>(s, fy * W + lxy + 3, a, aj, hy);
            oy = fy * TX + lxy;
            if constexpr (!HOLD) put_y(hy, oy);
        }
    }
    const int ic = (lyu + 3) * W + lxu + 3;  // the update cell's record: stage input U^(s-1)
    double qin[4] = {0, 0, 0, 0};
    if constexpr (HOLD) {
        if constexpr (ALIAS) {
            // the h arrays overwrite the records: take the update cell's input first
            if (vu) {
#pragma unroll
                for (int k = 0; k < 4; k++) qin[k] = s[k * NC + ic];
            }
        }
        __syncthreads();
        if (ox >= 0) put_x(hx, ox);
        if (oy >= 0) put_y(hy, oy);
    }
    __syncthreads();

    // ---- 3. RK stage update, time_advance_general.f90:35-54 / :113-135 / :195-217
    if (vu) {
        double q[4], qs[4];
#pragma unroll
        for (int k = 0; k < 4; k++) {
            const double dh_dxi =
                shx[k * T::NXF + lyu * (TX + 1) + lxu + 1] - shx[k * T::NXF + lyu * (TX + 1) + lxu];
            const double dh_deta = shy[k * T::NYF + (lyu + 1) * TX + lxu] - shy[k * T::NYF + lyu * TX + lxu];
            const double q1 = BANDS ? sc[k * TX * TY + lyu * TX + lxu] : ALIAS ? qin[k] : s[k * NC + ic];
            qs[k] = q1;
            if (a.stage == 1) {
                q[k] = q1 - dt / a.dx * dh_dxi - dt / a.dy * dh_deta;
            } else if (a.stage == 2) {
                q[k] = 3.0 / 4.0 * q0[k] + 1.0 / 4.0 * q1 - 1.0 / 4.0 * (dt / a.dx * dh_dxi + dt / a.dy * dh_deta);
            } else {
                q[k] = 1.0 / 3.0 * q0[k] + 2.0 / 3.0 * q1 - 2.0 / 3.0 * (dt / a.dx * dh_dxi + dt / a.dy * dh_deta);
            }
        }
        if (a.if_bodyforce) {
            // bodyforce = (0, 0, rho, rho v) of the stage input (boundary_rt_general.f90:154-155),
            // added in its own pass (time_advance_general.f90:71-84)
            const double coef = a.stage == 1 ? dt : (a.stage == 2 ? 1.0 / 4.0 * dt : TWO_THIRDS_SP * dt);
            q[2] = q[2] + coef * qs[0];
            q[3] = q[3] + coef * qs[2];
        }
#pragma unroll
        for (int k = 0; k < 4; k++) a.out[k * a.plane + ou] = q[k];
        double pr[4];
        prim_of(a.gamma, q, pr);
        const unsigned long long bi = dbits(fabs(pr[0]) + pr[3]), bj = dbits(fabs(pr[1]) + pr[3]);
        ei = bi > ei ? bi : ei;
        ej = bj > ej ? bj : ej;
    }
    __syncthreads();  // LDS is rewritten by the next tile
}

}  // namespace dense
