// Fused split-step kernel for the 2D relaxation-GRP Euler solver (gfx950).
//
// One launch advances every active instance by one full time step of the
// reference's scheme (src/wrappers.cuh:196-284):
//     A1 = X(U), A = Y(A1)      branch "X then Y"
//     B1 = Y(U), B = X(B1)      branch "Y then X"
//     U' = fill((A + B) / 2)    primitive average + ghost preset at t + dt
//     s_max(U') -> dt of the next step
// A wave owns a window of 64 consecutive columns (60 output columns plus a
// two-column halo on each side) and marches down a strip of R output rows.
// Both Y-sweeps are marched in registers along the rows (no transposes, no
// LDS); both X-sweeps exchange neighbour data across lanes with DPP
// wave_shr/wave_shl.  A1 and B1 never leave registers, so HBM sees one read
// of U (plus the halo rows) and one write of U'.
//
// The ghost preset (src/ghostCellPresets.cuh:47-241, two-phase row-then-
// column order) is applied by the lane that owns each ghost cell's source
// cell; constant inflow states are written by the clamped owner and enter
// the wave-speed maximum directly.  The maximum of the new state is reduced
// per wave (64-bit bit-pattern max, NaN sorts high), combined with one
// atomic per wave, and the last wave of an instance advances t, dt and the
// step counter on the device, so the host never waits inside the time loop.
#include <hip/hip_runtime.h>

#include <algorithm>

#include "fvm2d_device.h"
#include "fvm2d_kernels.h"
#include "fvm2d_ghost.h"
#include "fvm2d_lanes.h"
#include "fvm2d_march.h"

namespace fvm2d {

__device__ __forceinline__ unsigned long long dbits(double x) {
    return (unsigned long long)__double_as_longlong(x);
}

// State planes live in global memory: accessed through global-address-space
// pointers the loads and stores are global_* instructions, which count only
// on the vector memory counter (flat ones count on the LDS/scalar counter
// too, so every wait for an LDS exchange would also wait for the row loads
// in flight).
using gdouble = __attribute__((address_space(1))) double;

__device__ __forceinline__ void store4(double *base, long long plane, long long off, double r,
                                       double p, double u, double v) {
    gdouble *g = (gdouble *)base;
    g[off] = r;
    g[plane + off] = p;
    g[2 * plane + off] = u;
    g[3 * plane + off] = v;
}

// Write every ghost cell among the candidates whose fill source (or, for
// constants, whose clamped owner) is interior cell (j, c) holding value
// (r, p, u, v).  Candidates: rows {j, 0, 1, ny-2, ny-1} x columns
// {c, 0, 1, nx-2, nx-1} when corner != 0 (a cell in the first or last two
// interior rows and on the first or last interior column), else the
// ghost rows above and below column c (j in the first or last two
// interior rows) and the ghost columns beside row j (c on the first or
// last interior column).  Every ghost cell of the preset has its source
// or owner in exactly one of these sets.  Returns the largest signal-speed
// key among the written ghost cells; the caller folds it into its own
// maximum.  (Passed by reference, that maximum would live in private
// memory across the call, and its private-memory round trip made every
// row of the march wait for all of the previous row's stores.)
template <bool Exact>
__device__ __noinline__ unsigned long long emit_ghosts(const InstDesc &d, double *dst, int j, int c,
                                                       double r, double p, double u, double v,
                                                       double t_new, int corner) {
    unsigned long long mx = 0ull;
    const int nx = d.nx, ny = d.ny;
    const int rows[5] = {j, 0, 1, ny - 2, ny - 1};
    const int cols[5] = {c, 0, 1, nx - 2, nx - 1};
    const bool top = j == 2, bot = j >= ny - 4;
    const bool left = c == 2, right = c == nx - 3;
    for (int a = 0; a < 5; a++) {
        const int gj = rows[a];
        if (!corner && a > 0 && !((a <= 2 && top) || (a >= 3 && bot))) continue;
        for (int b = 0; b < 5; b++) {
            const int gi = cols[b];
            if (a == 0 && b == 0) continue;
            if (!corner && a > 0 && b > 0) continue;
            if (!corner && a == 0 && !((b <= 2 && left) || (b >= 3 && right))) continue;
            const bool ghost = gj < 2 || gj > ny - 3 || gi < 2 || gi > nx - 3;
            if (!ghost) continue;
            const GhostSrc g = ghost_source(d.strategy, gj, gi, nx, ny, t_new);
            const long long off = (long long)gj * d.pitch + gi;
            if (g.is_const) {
                if (clampi(gj, 2, ny - 3) == j && clampi(gi, 2, nx - 3) == c) {
                    store4(dst, d.plane, off, g.r, g.p, g.u, g.v);
                    const unsigned long long sb = dbits(speed<Exact>(g.r, g.p, g.u, g.v, d.gas));
                    mx = sb > mx ? sb : mx;
                }
            } else if (g.sj == j && g.si == c) {
                store4(dst, d.plane, off, r, p, u, g.negv ? -v : v);
            }
        }
    }
    return mx;
}

// Outflow preset (src/ghostCellPresets.cuh:106-145): every ghost cell is a
// copy of the nearest interior cell, so the owner of (j, c) writes it to
// the ghost cells beside it (c on the first/last interior column), above
// or below it (j on the first/last interior row), and to the corner
// blocks when both hold.
__device__ __forceinline__ void emit_outflow(const InstDesc &d, double *dst, int j, int c, double r,
                                             double p, double u, double v) {
    const int nx = d.nx, ny = d.ny;
    const long long P = d.plane, pitch = d.pitch;
    const int nr = (j == 2 ? 2 : 0) + (j == ny - 3 ? 2 : 0);
    const int nc = (c == 2 ? 2 : 0) + (c == nx - 3 ? 2 : 0);
    int gr[4], gc[4];
    int k = 0;
    if (j == 2) gr[k++] = 0, gr[k++] = 1;
    if (j == ny - 3) gr[k++] = ny - 2, gr[k++] = ny - 1;
    k = 0;
    if (c == 2) gc[k++] = 0, gc[k++] = 1;
    if (c == nx - 3) gc[k++] = nx - 2, gc[k++] = nx - 1;
    for (int a = 0; a < nr; a++) store4(dst, P, (long long)gr[a] * pitch + c, r, p, u, v);
    for (int b = 0; b < nc; b++) store4(dst, P, (long long)j * pitch + gc[b], r, p, u, v);
    for (int a = 0; a < nr; a++)
        for (int b = 0; b < nc; b++) store4(dst, P, (long long)gr[a] * pitch + gc[b], r, p, u, v);
}

__device__ __forceinline__ unsigned long long wave_max_u64(unsigned long long x) {
    for (int o = 32; o > 0; o >>= 1) {
        unsigned long long y = __shfl_xor(x, o, kWave);
        x = y > x ? y : x;
    }
    return x;
}

// dt rule of src/wrappers.cuh:202-207 evaluated on the device.
__device__ __forceinline__ void advance_ctl(const InstDesc &d, Ctl *c, double t_new, int steps_new,
                                            unsigned long long sbits) {
    const double smax = __longlong_as_double((long long)sbits);
    const double h = d.hmin;
    const double a = d.cfl * h / smax, b = d.t_end - t_new;
    const double dt = b < a ? b : a;  // std::min(a, b)
    int done = !(t_new < d.t_end);
    if (d.max_steps >= 0 && steps_new >= d.max_steps) done = 1;
    if (h / smax < 1e-20) done = 1;
    int status = 0;
    if (sbits > 0x7ff0000000000000ull) {  // NaN (either sign) sorts above +inf
        status = 1;
        done = 1;
    }
    c->dt_used = c->dt;
    c->t = t_new;
    c->dt = dt;
    c->smax = smax;
    c->steps = steps_new;
    c->done = done;
    c->status |= status;
    c->smax_next = 0ull;
}

// A control-block field read through the scalar cache.  The fields a task
// reads (t, dt, steps, done) change only when the instance's last task
// advances the block, after every other task of the launch has read them,
// and each launch starts with the scalar cache invalidated.
template <typename T>
__device__ __forceinline__ T ctl_field(const T *p) {
    return *(const __attribute__((address_space(4))) T *)p;
}

// ---------------------------------------------------------------- the step
// One task: window w (columns w * kOwn .. + kOwn - 1 owned), output rows
// [r0, r1).
template <bool Exact>
__device__ void step_task(const InstDesc &d, Ctl *ctl, int w, int r0, int r1, int lane, const Xch &xa,
                          const Xch &xb) {
    const int nx = d.nx, ny = d.ny;
    const long long P = d.plane;
    const double t = ctl_field(&ctl->t), dt = ctl_field(&ctl->dt);
    const int steps = ctl_field(&ctl->steps);
    const double *__restrict__ src = (steps & 1) ? d.buf1 : d.buf0;
    double *__restrict__ dst = (steps & 1) ? d.buf0 : d.buf1;
    const Gas G = d.gas;
    const Line LX = d.lx, LY = d.ly;
    const double lamx = dt / LX.h, lamy = dt / LY.h;
    const double t_new = t + dt;

    const int col = w * kOwn + lane;
    const int colc = col < nx ? col : nx - 1;
    const bool own = lane >= 2 && lane < kWave - 2 && col <= nx - 3;

    auto load = [&](int j) -> Q {  // X-order
        const long long o = (long long)j * d.pitch + colc;
        const gdouble *g = (const gdouble *)src;
        return Q{g[o], g[P + o], g[2 * P + o], g[3 * P + o]};
    };

    // prologue: rows r0-2 .. r0+1
    const Q u_m2 = load(r0 - 2), u_m1 = load(r0 - 1), u_0 = load(r0), u_p1 = load(r0 + 1);
    YMarch mu, ma;
    ymarch_init<Exact>(mu, swap_uv(u_m2), swap_uv(u_m1), swap_uv(u_0), swap_uv(u_p1), LY, G, dt);
    {
        const Q a_m2 = xsweep<Exact>(u_m2, LX, G, dt, lamx);
        const Q a_m1 = xsweep<Exact>(u_m1, LX, G, dt, lamx);
        const Q a_0 = xsweep<Exact>(u_0, LX, G, dt, lamx);
        const Q a_p1 = xsweep<Exact>(u_p1, LX, G, dt, lamx);
        ymarch_init<Exact>(ma, swap_uv(a_m2), swap_uv(a_m1), swap_uv(a_0), swap_uv(a_p1), LY, G,
                           dt);
    }

    unsigned long long mx = 0ull;
    Q u_n = load(r0 + 2);
    xput(xa, u_n);
    for (int j = r0; j < r1; j++) {
        const Q u_cur = u_n;
        const Q uL = xget(xa, xa.w), uR = xget(xa, xa.e);
        // prefetch the next row (unconditionally, clamped into the mesh: a
        // load on only some paths leaves the compiler waiting for it, and
        // for this row's stores, at the end of every iteration)
        u_n = load(j + 3 < ny ? j + 3 : ny - 1);
        // branch B, first sweep: B1 row j (exchanged at once, needed after A)
        const Q b1 = swap_uv(ymarch_step<Exact>(mu, swap_uv(u_cur), LY, G, dt, lamy));
        xput(xb, b1);
        // branch A, first sweep on the look-ahead row: A1 row j+2
        const Q a1 = xsweep_pre<Exact>(u_cur, uL, uR, LX, G, dt, lamx);
        // B1's neighbours, read a Y-march step ahead of their use
        const Q bL = xget(xb, xb.w), bR = xget(xb, xb.e);
        // branch A, second sweep: A row j
        const Q a = swap_uv(ymarch_step<Exact>(ma, swap_uv(a1), LY, G, dt, lamy));
        // branch B, second sweep: B row j
        const Q b = xsweep_pre<Exact>(b1, bL, bR, LX, G, dt, lamx);
        // the next row's exchange, an iteration ahead of its use
        xput(xa, u_n);
        // src/basicFuns.cuh:264-274 average on primitives
        const double r = (a.r + b.r) * 0.5, p = (a.p + b.p) * 0.5;
        const double u = (a.un + b.un) * 0.5, v = (a.ut + b.ut) * 0.5;
        if (own) {
            const long long o = (long long)j * d.pitch + col;
            store4(dst, P, o, r, p, u, v);
            const unsigned long long sb = dbits(speed<Exact>(r, p, u, v, G));
            mx = sb > mx ? sb : mx;
            const bool edge_row = j == 2 || j >= ny - 4, edge_col = col == 2 || col == nx - 3;
            if (edge_row || edge_col) {
                if (d.strategy == kOutflow)
                    emit_outflow(d, dst, j, col, r, p, u, v);
                else
                {
                    const unsigned long long gm =
                        emit_ghosts<Exact>(d, dst, j, col, r, p, u, v, t_new, edge_row && edge_col);
                    mx = gm > mx ? gm : mx;
                }
            }
        }
    }

    // The wave's speed maximum goes into the instance's accumulator with one
    // device-scope atomic whose result is not waited for; the advance kernel
    // launched after this one reads the total.  (An earlier build let the
    // last task of the launch advance the control block: its arrival
    // counters needed a device-scope fence per wave, which on this
    // many-XCD part writes back and invalidates the XCD's L2 under the
    // other waves' feet: 2000 x 2000 ran 181 instead of 150 us per step.)
    mx = wave_max_u64(mx);
    if (lane == 0) atomicMax(&ctl->smax_next, mx);
}

template <bool Exact>
__global__ void __launch_bounds__(256) fvm_step_kernel(const InstDesc *__restrict__ descs,
                                                       Ctl *__restrict__ ctls,
                                                       const int4 *__restrict__ tasks, int ntasks,
                                                       Sched *sched, int persistent) {
    const int lane = threadIdx.x & (kWave - 1);
    const int wid = threadIdx.x / kWave;
    const int wpb = blockDim.x / kWave;
    // per wave: the current input row's and the B1 row's exchange buffers
    __shared__ double xbuf[4][2][4 * kWave];
    const Xch xa = make_xch(xbuf[wid][0], lane), xb = make_xch(xbuf[wid][1], lane);
    int tk;
    if (persistent) {
        unsigned int v = 0;
        if (lane == 0) v = atomicAdd(&sched->ticket, 1u);
        tk = __builtin_amdgcn_readfirstlane((int)__shfl(v, 0, kWave));
    } else {
        tk = blockIdx.x * wpb + wid;
    }
    while (tk < ntasks) {
        const int4 task = tasks[tk];
        const int inst = __builtin_amdgcn_readfirstlane(task.x);
        const int w = __builtin_amdgcn_readfirstlane(task.y);
        const int r0 = __builtin_amdgcn_readfirstlane(task.z);
        const int r1 = __builtin_amdgcn_readfirstlane(task.w);
        Ctl *c = ctls + inst;
        const InstDesc &d = descs[inst];
        if (!ctl_field(&c->done)) step_task<Exact>(d, c, w, r0, r1, lane, xa, xb);
        if (!persistent) break;
        unsigned int v = 0;
        if (lane == 0) v = atomicAdd(&sched->ticket, 1u);
        tk = __builtin_amdgcn_readfirstlane((int)__shfl(v, 0, kWave));
    }
}

// After every step launch: each instance still running takes its step --
// t += dt, the step count, the dt rule on the new state's s_max (the
// accumulator the step's waves filled) -- and the persistent launch's task
// ticket is reset.  One thread per instance.
__global__ void fvm_advance_kernel(const InstDesc *__restrict__ descs, Ctl *__restrict__ ctls, int ninst,
                                   Sched *sched) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i == 0 && sched) sched->ticket = 0u;
    if (i >= ninst) return;
    Ctl *c = ctls + i;
    if (c->done) return;
    advance_ctl(descs[i], c, c->t + c->dt, c->steps + 1, c->smax_next);
}

// ---------------------------------------------------------------- set-up
// s_max over every cell of the current state (ghosts included), as
// maxPropagtingSpeed + thrust::reduce do at src/wrappers.cuh:198-201.  Blocks
// stride over rows, threads over columns; one atomic per block.
template <bool Exact>
__global__ void __launch_bounds__(256) fvm_speed_kernel(const InstDesc *__restrict__ descs,
                                                        Ctl *__restrict__ ctls, int inst,
                                                        int use_buf1) {
    __shared__ unsigned long long red[256 / kWave];
    const InstDesc d = descs[inst];
    const double *src = use_buf1 ? d.buf1 : d.buf0;
    unsigned long long mx = 0ull;
    for (int j = blockIdx.x; j < d.ny; j += gridDim.x) {
        for (int i = threadIdx.x; i < d.nx; i += blockDim.x) {
            const long long o = (long long)j * d.pitch + i;
            const double sp = speed<Exact>(src[o], src[d.plane + o], src[2 * d.plane + o],
                                           src[3 * d.plane + o], d.gas);
            const unsigned long long b = dbits(sp);
            mx = b > mx ? b : mx;
        }
    }
    mx = wave_max_u64(mx);
    const int wv = threadIdx.x / kWave;
    if ((threadIdx.x & (kWave - 1)) == 0) red[wv] = mx;
    __syncthreads();
    if (threadIdx.x == 0) {
        for (int k = 1; k < (int)(blockDim.x / kWave); k++) mx = red[k] > mx ? red[k] : mx;
        atomicMax(&ctls[inst].smax_next, mx);
    }
}

__global__ void fvm_ctl_start_kernel(const InstDesc *__restrict__ descs, Ctl *__restrict__ ctls,
                                     int ninst, const double *__restrict__ t0s) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= ninst) return;
    Ctl *c = ctls + i;
    const unsigned long long sb = c->smax_next;
    // the reference's loop test happens before the first step (t = t0 = 0)
    const InstDesc d = descs[i];
    c->status = 0;
    advance_ctl(d, c, t0s ? t0s[i] : 0.0, 0, sb);
}

// Resume time loops from saved control blocks (the state that goes with
// each was saved beside it and is copied to buf0): t, dt and s_max as they
// were, the step count restarted at 0 so that buf0 is the current buffer.
__global__ void fvm_ctl_restore_kernel(Ctl *__restrict__ ctls, const Ctl *__restrict__ saved,
                                       int ninst) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= ninst) return;
    Ctl c = saved[i];
    c.steps = 0;
    c.smax_next = 0ull;
    ctls[i] = c;
}

// ---------------------------------------------------------------- record frames
// One cell of a 1-D sweep from the five cells of its line around it (line
// order: m2, m1 are the "left" neighbours, p1, p2 the "right" ones); the same
// operations as the lane-parallel sweeps.
template <bool Exact>
__device__ Q line_update(const Q &m2, const Q &m1, const Q &q0, const Q &p1, const Q &p2,
                         const Line &L, const Gas &G, double dt) {
    const Q s_m1 = sweep_slope<Exact>(m2, m1, q0, L);
    const Q s_0 = sweep_slope<Exact>(m1, q0, p1, L);
    const Q s_p1 = sweep_slope<Exact>(q0, p1, p2, L);
    const double k = sweep_k<Exact>(dt, L);
    IfState W, Wt;
    grp<Exact>(sweep_face_plus<Exact>(m1, s_m1, L), sweep_face_minus<Exact>(q0, s_0, L), s_m1, s_0, G,
               W, Wt);
    const Q FL = flux<Exact>(W, Wt, k);
    grp<Exact>(sweep_face_plus<Exact>(q0, s_0, L), sweep_face_minus<Exact>(p1, s_p1, L), s_0, s_p1, G,
               W, Wt);
    const Q FR = flux<Exact>(W, Wt, k);
    return update<Exact>(q0, FL, FR, dt / L.h, G);
}

// The state after a step as the reference's record mode copies it out:
// after the branch average, before the ghost fill (src/wrappers.cuh:248-261,
// fill at :265-281).  From the step's input U (prev) and output (cur):
//   interior cells          cur (the fill does not touch them);
//   ghost rows, columns 2..nx-3: X(U) -- the X-sweep updates every row, the
//       Y-sweep leaves rows 0, 1, ny-2, ny-1 alone, so both branches hold
//       X(U) there and their average is X(U) exactly;
//   ghost columns, rows 2..ny-3: Y(U), by the same argument;
//   the four 2 x 2 corners: U (neither sweep touches them).
// out: four planes in the state's padded layout.
template <bool Exact>
__global__ void __launch_bounds__(256) fvm_frame_kernel(InstDesc d, const double *__restrict__ prev,
                                                        const double *__restrict__ cur, double dt,
                                                        double *__restrict__ out) {
    const long long n = (long long)d.nx * d.ny;
    const long long P = d.plane;
    for (long long k = (long long)blockIdx.x * blockDim.x + threadIdx.x; k < n;
         k += (long long)gridDim.x * blockDim.x) {
        const int j = (int)(k / d.nx), i = (int)(k % d.nx);
        const long long o = (long long)j * d.pitch + i;
        const bool grow = j < 2 || j > d.ny - 3, gcol = i < 2 || i > d.nx - 3;
        auto at = [&](int jj, int ii) {
            const long long q = (long long)jj * d.pitch + ii;
            return Q{prev[q], prev[P + q], prev[2 * P + q], prev[3 * P + q]};
        };
        Q v;
        if (!grow && !gcol) {
            v = Q{cur[o], cur[P + o], cur[2 * P + o], cur[3 * P + o]};
        } else if (grow && gcol) {
            v = at(j, i);
        } else if (grow) {
            v = line_update<Exact>(at(j, i - 2), at(j, i - 1), at(j, i), at(j, i + 1), at(j, i + 2), d.lx,
                                   d.gas, dt);
        } else {
            // Y line order: the left neighbour of row j is row j + 1, u and v swapped
            v = swap_uv(line_update<Exact>(swap_uv(at(j + 2, i)), swap_uv(at(j + 1, i)),
                                           swap_uv(at(j, i)), swap_uv(at(j - 1, i)),
                                           swap_uv(at(j - 2, i)), d.ly, d.gas, dt));
        }
        out[o] = v.r;
        out[P + o] = v.p;
        out[2 * P + o] = v.un;
        out[3 * P + o] = v.ut;
    }
}

void launch_frame(bool exact, const InstDesc &d, const double *prev, const double *cur, double dt,
                  double *out, hipStream_t stream) {
    const long long n = (long long)d.nx * d.ny;
    const unsigned blocks = (unsigned)std::min<long long>((n + 255) / 256, 4096);
    if (exact)
        hipLaunchKernelGGL(fvm_frame_kernel<true>, dim3(blocks), dim3(256), 0, stream, d, prev, cur, dt, out);
    else
        hipLaunchKernelGGL(fvm_frame_kernel<false>, dim3(blocks), dim3(256), 0, stream, d, prev, cur, dt, out);
}

// ---------------------------------------------------------------- 8-bit frame image
// The cv::normalize(..., 0, 255, NORM_MINMAX, CV_8UC1) step of the
// reference's frame writer (src/wrappers.cuh:323-340) on a frame's density:
// lo and hi over every cell (ghosts included), s = 255 / (hi - lo), pixel =
// rint(rho s - lo s) clamped to 0..255; every pixel 0 when hi == lo.  NaN
// densities are left out of lo and hi and give 0.
//
// Order-preserving 64-bit keys of doubles (non-NaN): min and max are exact,
// so atomics in any order give the same lo and hi.
__device__ __forceinline__ unsigned long long order_key(double x) {
    const unsigned long long b = dbits(x);
    return (b >> 63) ? ~b : (b | 0x8000000000000000ull);
}

__device__ __forceinline__ double key_value(unsigned long long k) {
    return __longlong_as_double((long long)((k >> 63) ? (k & 0x7fffffffffffffffull) : ~k));
}

__global__ void fvm_range_init_kernel(unsigned long long *range) {
    if (threadIdx.x == 0) {
        range[0] = ~0ull;  // min key
        range[1] = 0ull;   // max key
    }
}

__global__ void __launch_bounds__(256) fvm_range_kernel(InstDesc d, const double *__restrict__ frame,
                                                        unsigned long long *range) {
    unsigned long long lo = ~0ull, hi = 0ull;
    for (int j = blockIdx.x; j < d.ny; j += gridDim.x)
        for (int i = threadIdx.x; i < d.nx; i += blockDim.x) {
            const double r = frame[(long long)j * d.pitch + i];
            if (r != r) continue;
            const unsigned long long key = order_key(r);
            lo = key < lo ? key : lo;
            hi = key > hi ? key : hi;
        }
    lo = ~wave_max_u64(~lo);
    hi = wave_max_u64(hi);
    if ((threadIdx.x & (kWave - 1)) == 0) {
        atomicMin(&range[0], lo);
        atomicMax(&range[1], hi);
    }
}

__global__ void __launch_bounds__(256) fvm_gray_kernel(InstDesc d, const double *__restrict__ frame,
                                                       const unsigned long long *__restrict__ range,
                                                       unsigned char *__restrict__ gray) {
    const unsigned long long klo = range[0], khi = range[1];
    const bool any = klo <= khi;
    const double lo = any ? key_value(klo) : 0.0, hi = any ? key_value(khi) : 0.0;
    const double s = hi > lo ? 255.0 / (hi - lo) : 0.0;
    const double ls = lo * s;
    for (int j = blockIdx.x; j < d.ny; j += gridDim.x)
        for (int i = threadIdx.x; i < d.nx; i += blockDim.x) {
            double v = rint(frame[(long long)j * d.pitch + i] * s - ls);
            v = v > 0.0 ? v : 0.0;  // NaN -> 0
            v = v < 255.0 ? v : 255.0;
            gray[(long long)j * d.nx + i] = (unsigned char)v;
        }
}

void launch_frame_gray(const InstDesc &d, const double *frame, unsigned long long *range,
                       unsigned char *gray, hipStream_t stream) {
    const unsigned blocks = (unsigned)std::min(d.ny, 4096);
    hipLaunchKernelGGL(fvm_range_init_kernel, dim3(1), dim3(64), 0, stream, range);
    hipLaunchKernelGGL(fvm_range_kernel, dim3(blocks), dim3(256), 0, stream, d, frame, range);
    hipLaunchKernelGGL(fvm_gray_kernel, dim3(blocks), dim3(256), 0, stream, d, frame, range, gray);
}

// ---------------------------------------------------------------- launchers
void launch_step(bool exact, const InstDesc *descs, Ctl *ctls, int ninst, const int4 *tasks, int ntasks,
                 Sched *sched, int persistent, int grid, int wpb, hipStream_t stream) {
    const dim3 block(wpb * kWave);
    if (exact)
        hipLaunchKernelGGL(fvm_step_kernel<true>, dim3(grid), block, 0, stream, descs, ctls, tasks,
                           ntasks, sched, persistent);
    else
        hipLaunchKernelGGL(fvm_step_kernel<false>, dim3(grid), block, 0, stream, descs, ctls, tasks,
                           ntasks, sched, persistent);
    hipLaunchKernelGGL(fvm_advance_kernel, dim3((ninst + 255) / 256), dim3(256), 0, stream, descs, ctls,
                       ninst, persistent ? sched : nullptr);
}

void launch_speed(bool exact, const InstDesc *descs, Ctl *ctls, int inst, int use_buf1,
                  long long nrows, hipStream_t stream) {
    long long blocks = nrows;
    if (blocks > 512) blocks = 512;
    if (blocks < 1) blocks = 1;
    if (exact)
        hipLaunchKernelGGL(fvm_speed_kernel<true>, dim3((unsigned)blocks), dim3(256), 0, stream,
                           descs, ctls, inst, use_buf1);
    else
        hipLaunchKernelGGL(fvm_speed_kernel<false>, dim3((unsigned)blocks), dim3(256), 0, stream,
                           descs, ctls, inst, use_buf1);
}

void launch_ctl_start(const InstDesc *descs, Ctl *ctls, int ninst, const double *t0s,
                      hipStream_t stream) {
    hipLaunchKernelGGL(fvm_ctl_start_kernel, dim3((ninst + 63) / 64), dim3(64), 0, stream, descs,
                       ctls, ninst, t0s);
}

void launch_ctl_restore(Ctl *ctls, const Ctl *saved, int ninst, hipStream_t stream) {
    hipLaunchKernelGGL(fvm_ctl_restore_kernel, dim3((ninst + 63) / 64), dim3(64), 0, stream, ctls,
                       saved, ninst);
}

int step_kernel_occupancy(bool exact, int wpb) {
    int blocks = 0;
    hipError_t e;
    if (exact)
        e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&blocks, fvm_step_kernel<true>,
                                                         wpb * kWave, 0);
    else
        e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&blocks, fvm_step_kernel<false>,
                                                         wpb * kWave, 0);
    return e == hipSuccess ? blocks : 0;
}

}  // namespace fvm2d
