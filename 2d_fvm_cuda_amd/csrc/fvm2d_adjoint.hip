// Reverse (adjoint) step of the split scheme on gfx950.
//
// Forward step (fixed by the state U_n):  dt = CFL h / s_max(U_n)
//     A1 = X(U), A = Y(A1);  B1 = Y(U), B = X(B1);  U' = fill((A + B)/2)
// Reverse step, given lam' = dL/dU' (primitive planes, every cell):
//   R1  recompute A1 = X(U), B1 = Y(U); s_max(U) and dt come from the forward
//       time loop's control block of U, and the row sweeps find the first
//       cell attaining s_max (its arg-max)
//   R2  lam_A1 = Y^T(A1; lam_C/2),  lam_B1 = X^T(B1; lam_C/2)
//       where lam_C = fill^T(lam') is gathered while loading
//   R3  lam = Y^T(U; lam_B1) + X^T(U; lam_A1)
//   R4  lam[argmax] += dtbar * d dt / d U   (dt depends on the arg-max cell)
// Every sweep adjoint is one lane-parallel routine: lanes are 64
// consecutive cells of one line; an X line is a row (unit stride), a Y line
// is a column read with the row pitch as stride, in the orientation of the
// reference's rotated sweep (line index k = ny-1-j) and with u and v
// swapped.  The routine recomputes the sweep's forward intermediates (limited
// slopes, faces, GRP states, fluxes, updated states) and applies the
// hand-written adjoints of fvm2d_adjoint_cell.h; a cell's adjoint depends on
// cells up to four away, so 56 of the 64 lanes produce output.
#include <hip/hip_runtime.h>

#include <algorithm>

#include "fvm2d_adjoint_cell.h"
#include "fvm2d_adjoint_kernels.h"
#include "fvm2d_ghost.h"
#include "fvm2d_lanes.h"
#include "fvm2d_march.h"

namespace fvm2d {

constexpr int kAdjHalo = 4;
constexpr int kAdjOwn = kWave - 2 * kAdjHalo;  // 56
constexpr int kFwdHalo = 2;
constexpr int kFwdOwn = kWave - 2 * kFwdHalo;  // 60

// State and adjoint planes live in global memory: accessed through
// global-address-space pointers the loads and stores are global_*
// instructions, which count only on the vector memory counter (flat ones
// also count on the counter that DPP/LDS and scalar-load waits use).
using gdouble = __attribute__((address_space(1))) double;
__device__ __forceinline__ const gdouble *gp(const double *a) { return (const gdouble *)a; }
__device__ __forceinline__ gdouble *gp(double *a) { return (gdouble *)a; }

// Geometry of one lane line of a task.
struct LineMap {
    long long base;    // offset of line index 0
    long long stride;  // offset step per line index (+1 for X, -pitch for Y)
    int n;             // cells on the line
    int k;             // this lane's line index
    int kc;            // clamped into [0, n)
    bool ydir;
};

__device__ __forceinline__ LineMap line_map(const AdjDesc &d, const int4 &t, int lane, int halo,
                                            int own) {
    LineMap m;
    m.ydir = t.y != 0;
    m.k = t.z * own - halo + lane;
    if (!m.ydir) {  // row t.w, line index = column
        m.n = d.nx;
        m.base = (long long)t.w * d.pitch;
        m.stride = 1;
    } else {  // column t.w, line index k = ny-1-row
        m.n = d.ny;
        m.base = (long long)(d.ny - 1) * d.pitch + t.w;
        m.stride = -(long long)d.pitch;
    }
    m.kc = m.k < 0 ? 0 : (m.k >= m.n ? m.n - 1 : m.k);
    return m;
}

__device__ __forceinline__ long long off_of(const LineMap &m) { return m.base + m.kc * m.stride; }

__device__ __forceinline__ Q load_line(const double *ap, long long plane, const LineMap &m) {
    const long long o = off_of(m);
    const gdouble *a = gp(ap);
    const Q q{a[o], a[plane + o], a[2 * plane + o], a[3 * plane + o]};
    return pick(m.ydir, swap_uv(q), q);
}

__device__ __forceinline__ void store_line(double *ap, long long plane, const LineMap &m, const Q &q,
                                           bool accumulate) {
    const long long o = off_of(m);
    gdouble *a = gp(ap);
    const Q x = pick(m.ydir, swap_uv(q), q);
    if (accumulate) {
        a[o] += x.r;
        a[plane + o] += x.p;
        a[2 * plane + o] += x.un;
        a[3 * plane + o] += x.ut;
    } else {
        a[o] = x.r;
        a[plane + o] = x.p;
        a[2 * plane + o] = x.un;
        a[3 * plane + o] = x.ut;
    }
}

// Geometry of the task's line, selected field by field (a selected struct
// reference would be materialised in private memory).
__device__ __forceinline__ Line line_of(const AdjDesc &d, const LineMap &m) {
    Line L;
    L.h = m.ydir ? d.ly.h : d.lx.h;
    L.ih = m.ydir ? d.ly.ih : d.lx.ih;
    L.ih_a = m.ydir ? d.ly.ih_a : d.lx.ih_a;
    L.ih_c = m.ydir ? d.ly.ih_c : d.lx.ih_c;
    L.hh = m.ydir ? d.ly.hh : d.lx.hh;
    L.alpha = m.ydir ? d.ly.alpha : d.lx.alpha;
    L.quarter = d.lx.quarter;
    return L;
}

__device__ __forceinline__ unsigned long long bits_of(double x) {
    return (unsigned long long)__double_as_longlong(x);
}

__device__ __forceinline__ double wave_sum(double x) {
    for (int o = 32; o > 0; o >>= 1) x += __shfl_xor(x, o, kWave);
    return x;
}

// fill^T: adjoint of the filled state U' at interior cell (j, c) (zero for
// ghost cells: the fill overwrites them), plus the adjoints of every ghost
// cell whose fill copies (j, c).  Split in the plain load at the cell
// (issued ahead with the task's other loads) and the ghost additions, which
// only cells on the first/last two interior rows or columns have.
__device__ __forceinline__ bool is_interior(const AdjDesc &d, int j, int c) {
    return !(j < 2 || j > d.ny - 3 || c < 2 || c > d.nx - 3);
}

__device__ __forceinline__ Q lam_at(const AdjDesc &d, const double *lamp, int j, int c) {
    const long long P = d.plane, o = (long long)j * d.pitch + c;
    const gdouble *lam = gp(lamp);
    const Q s{lam[o], lam[P + o], lam[2 * P + o], lam[3 * P + o]};
    const bool in = is_interior(d, j, c);
    return Q{in ? s.r : 0.0, in ? s.p : 0.0, in ? s.un : 0.0, in ? s.ut : 0.0};
}

__device__ Q lamC_ghosts(const AdjDesc &d, const double *lamp, int j, int c, Q s) {
    const int nx = d.nx, ny = d.ny;
    const gdouble *lam = gp(lamp);
    if (!is_interior(d, j, c)) return s;
    if (j == 2 || j == ny - 4 || j == ny - 3 || c == 2 || c == nx - 3) {
        const long long P = d.plane;
        const int rows[5] = {j, 0, 1, ny - 2, ny - 1};
        const int cols[5] = {c, 0, 1, nx - 2, nx - 1};
#pragma unroll
        for (int a = 0; a < 5; a++) {
#pragma unroll
            for (int b = 0; b < 5; b++) {
                if (a == 0 && b == 0) continue;
                const int gj = rows[a], gi = cols[b];
                if (!(gj < 2 || gj > ny - 3 || gi < 2 || gi > nx - 3)) continue;
                const GhostSrc g = ghost_source(d.strategy, gj, gi, nx, ny, 0.0);
                if (g.is_const || g.sj != j || g.si != c) continue;
                const long long o = (long long)gj * d.pitch + gi;
                s.r += lam[o];
                s.p += lam[P + o];
                s.un += lam[2 * P + o];
                s.ut += g.negv ? -lam[3 * P + o] : lam[3 * P + o];
            }
        }
    }
    return s;
}

// Workgroups are dispatched to the 8 XCDs round robin (block b on XCD
// b mod 8), and each XCD has its own L2.  Tasks of neighbouring columns read
// the same cache lines (16 columns per 128-byte line), so a grid of a
// multiple of 8 blocks is renumbered to give each XCD a contiguous range of
// block indices, hence of tasks, instead of every 8th.
constexpr int kXcds = 8;
__device__ __forceinline__ int xcd_block() {
    const int b = blockIdx.x, n = gridDim.x;
    if (n % kXcds) return b;
    return (b % kXcds) * (n / kXcds) + b / kXcds;
}

// A value of the forward loop's control block (constant while the reverse
// step runs), read through the scalar cache: a vector load would make the
// task wait on the vector memory counter, behind its look-ahead loads.
__device__ __forceinline__ double ctl_value(const double *p) {
    return *(const __attribute__((address_space(4))) double *)p;
}

// The task a wave works on is wave-uniform: keep it in scalar registers.
__device__ __forceinline__ int4 uniform_task(const int4 *__restrict__ tasks, int tk) {
    const int4 v = tasks[tk];
    return make_int4(__builtin_amdgcn_readfirstlane(v.x), __builtin_amdgcn_readfirstlane(v.y),
                     __builtin_amdgcn_readfirstlane(v.z), __builtin_amdgcn_readfirstlane(v.w));
}

// ---------------------------------------------------------------- R1
// Forward first sweeps A1 = X(U) (X tasks) and B1 = Y(U) (Y tasks).  Each
// wave walks a strided list of tasks and loads the next task's line while it
// computes the current one.
__global__ void __launch_bounds__(256) adj_fwd_kernel(AdjDesc *__restrict__ descs,
                                                      const int4 *__restrict__ tasks, int ntasks) {
    const int lane = threadIdx.x & (kWave - 1);
    const int nwv = gridDim.x * (blockDim.x / kWave);
    int tk = __builtin_amdgcn_readfirstlane(blockIdx.x * (blockDim.x / kWave) + threadIdx.x / kWave);
    if (tk >= ntasks) return;
    int4 t = uniform_task(tasks, tk);
    Q q = load_line(descs[t.x].U, descs[t.x].plane, line_map(descs[t.x], t, lane, kFwdHalo, kFwdOwn));
    for (;;) {
        const int tn = tk + nwv;
        const bool more = tn < ntasks;
        int4 t2 = t;
        Q q2 = q;
        if (more) {
            t2 = uniform_task(tasks, tn);
            q2 = load_line(descs[t2.x].U, descs[t2.x].plane, line_map(descs[t2.x], t2, lane, kFwdHalo, kFwdOwn));
        }
        // the line geometry is recomputed from the task, not carried over
        // (a carried struct copy went through memory)
        const LineMap m = line_map(descs[t.x], t, lane, kFwdHalo, kFwdOwn);
        const AdjDesc &d = descs[t.x];
        const Line L = line_of(d, m);
        const double dt = ctl_value(&d.fctl->dt);
        const Q qo = xsweep<false>(q, L, d.gas, dt, dt / L.h);
        const bool own = lane >= kFwdHalo && lane < kWave - kFwdHalo && m.k < m.n;
        const bool upd = m.k >= 2 && m.k <= m.n - 3;
        if (own) store_line(m.ydir ? d.B1 : d.A1, d.plane, m, pick(upd, qo, q), false);
        if (!m.ydir) {
            // the X tasks cover every cell of U once: the first cell whose
            // wave speed is s_max (the forward loop's maximum over all cells,
            // by the same function) is the arg-max.  Regions of one constant
            // state hold many such cells: only the wave's first (its lowest
            // index) competes, and only while it is below the minimum found
            // so far, so that the atomics do not queue on the one address.
            const unsigned long long sb = bits_of(ctl_value(&d.fctl->smax));
            const bool hit = own && bits_of(speed<false>(q.r, q.p, q.un, q.ut, d.gas)) == sb;
            const long long idx = (long long)t.w * d.nx + m.k;
            const unsigned long long hits = __ballot(hit);
            if (hits != 0ull) {
                const int first = (int)__builtin_ctzll(hits);
                if (lane == first &&
                    idx < __hip_atomic_load(&d.ctl->argmax, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT))
                    atomicMin(&d.ctl->argmax, idx);
            }
            if (lane == 0 && !(sb <= 0x7ff0000000000000ull)) atomicOr(&d.ctl->status, 1);
        }
        if (!more) break;
        tk = tn;
        t = t2;
        q = q2;
    }
}

// Forward first sweep B1 = Y(U) with the step kernel's register march: a
// wave owns 64 whole columns (the Y-sweep is column-local, no halo) and a
// strip of interior rows [r0, r1) (task {instance, 64-column window, r0,
// r1}); it reads rows r0-2 .. r1+1 along the columns, so every load and
// store is a row segment.  Rows 0, 1, ny-2 and ny-1 are not swept (the
// reference updates lines [2, n-3]): the first and last strips copy them
// from U.  Same operations on the same values as the line form above.
__global__ void __launch_bounds__(256) adj_fwd_y_kernel(AdjDesc *__restrict__ descs,
                                                        const int4 *__restrict__ tasks, int ntasks) {
    const int lane = threadIdx.x & (kWave - 1);
    const int tk = __builtin_amdgcn_readfirstlane(blockIdx.x * (blockDim.x / kWave) + threadIdx.x / kWave);
    if (tk >= ntasks) return;
    const int4 t = uniform_task(tasks, tk);
    const AdjDesc &d = descs[t.x];
    const int nx = d.nx, ny = d.ny, r0 = t.z, r1 = t.w;
    const long long P = d.plane, pitch = d.pitch;
    const gdouble *__restrict__ U = gp(d.U);
    gdouble *__restrict__ B1 = gp(d.B1);
    const int col = t.y * kWave + lane;
    const int colc = col < nx ? col : nx - 1;
    const bool valid = col < nx;
    const double dt = ctl_value(&d.fctl->dt);
    const Line LY = d.ly;
    const double lamy = dt / LY.h;
    auto load = [&](int j) -> Q {
        const long long o = (long long)j * pitch + colc;
        return Q{U[o], U[P + o], U[2 * P + o], U[3 * P + o]};
    };
    auto store = [&](int j, const Q &q) {
        if (!valid) return;
        const long long o = (long long)j * pitch + col;
        B1[o] = q.r;
        B1[P + o] = q.p;
        B1[2 * P + o] = q.un;
        B1[3 * P + o] = q.ut;
    };
    if (r0 == 2) {
        store(0, load(0));
        store(1, load(1));
    }
    if (r1 == ny - 2) {
        store(ny - 2, load(ny - 2));
        store(ny - 1, load(ny - 1));
    }
    YMarch mu;
    ymarch_init<false>(mu, swap_uv(load(r0 - 2)), swap_uv(load(r0 - 1)), swap_uv(load(r0)),
                       swap_uv(load(r0 + 1)), LY, d.gas, dt);
    Q u_n = load(r0 + 2);
    for (int j = r0; j < r1; j++) {
        const Q u_cur = u_n;
        if (j + 1 < r1) u_n = load(j + 3);
        store(j, swap_uv(ymarch_step<false>(mu, swap_uv(u_cur), LY, d.gas, dt, lamy)));
    }
}

// ---------------------------------------------------------------- R2 / R3
// One sweep adjoint per task.  mode 0 (R2): input state A1 (Y tasks) or B1
// (X tasks), output adjoint lam_C/2 gathered from lam_in, result written to
// lam_A1 / lam_B1.  mode 2 (R3b, Y tasks, launched first): input U, output
// adjoint lam_B1, result written to lam_out.  mode 1 (R3a, X tasks): input
// U, output adjoint lam_A1, result added to lam_out -- the accumulation's
// extra load and store are row segments there, not column-strided.

// park: the wave's 4 x 64 doubles of LDS.
__device__ __forceinline__ void sweep_task(const AdjDesc &d, const LineMap &m, int lane, int mode, int part,
                                           double *park) {
    const Line L = line_of(d, m);
    const Gas &G = d.gas;
    const double dt = ctl_value(&d.fctl->dt);
    const double lam = dt / L.h;
    const long long P = d.plane;
    const double *src = mode == 0 ? (m.ydir ? d.A1 : d.B1) : d.U;

    const bool inl = m.k >= 0 && m.k < m.n;
    const bool upd = inl && m.k >= 2 && m.k <= m.n - 3;
    const bool sig = inl && m.k >= 1 && m.k <= m.n - 2;
    const bool ifc = inl && m.k >= 2 && m.k <= m.n - 2;
    const bool own = inl && lane >= kAdjHalo && lane < kWave - kAdjHalo;

    // ---- forward recompute of the sweep
    Q fpL, fm, sL, s;
    IfState W, Wt;
    Q F, FR, qo;
    {
        const Q q = load_line(src, P, m);
        const Q qL = shr(q), qR = shl(q);
        s = slope<false>(qL, q, qR, L);
        fm = face_minus<false>(q, s, L);
        const Q fp = face_plus<false>(q, s, L);
        fpL = shr(fp);
        sL = shr(s);
        grp<false>(fpL, fm, sL, s, G, W, Wt);
        F = flux<false>(W, Wt, dt);
        FR = shl(F);
        qo = update<false>(q, F, FR, lam, G);
    }
    // ---- reverse: cell update and flux
    Q qbo;
    {
        const long long o = off_of(m);
        if (mode == 0) {
            const Q s0 = lamC_ghosts(d, d.lam_in, (int)(o / d.pitch), (int)(o % d.pitch),
                                     lam_at(d, d.lam_in, (int)(o / d.pitch), (int)(o % d.pitch)));
            const Q h{0.5 * s0.r, 0.5 * s0.p, 0.5 * s0.un, 0.5 * s0.ut};
            qbo = pick(m.ydir, swap_uv(h), h);
        } else {
            qbo = load_line(m.ydir ? d.lamB1 : d.lamA1, P, m);
        }
    }
    double dtb = 0.0;
    Q Ub{0.0, 0.0, 0.0, 0.0};
    if (upd) {
        Ub = prim_adj(qo, qbo, G);
        dtb -= ((FR.r - F.r) * Ub.r + (FR.p - F.p) * Ub.p + (FR.un - F.un) * Ub.un +
                (FR.ut - F.ut) * Ub.ut) * L.ih;
    }
    const Q UbL = shr(Ub);
    // the cell's own term needs Ub (updated cells) or the output adjoint
    // (copied cells) after the GRP adjoint: parked in wave-private LDS across
    // it instead of held in four registers (the column sweep 3.7 % faster)
    {
        const Q x = pick(upd, Ub, qbo);
        park[lane] = x.r;
        park[kWave + lane] = x.p;
        park[2 * kWave + lane] = x.un;
        park[3 * kWave + lane] = x.ut;
    }
    Q fpLb{0.0, 0.0, 0.0, 0.0}, fmb{0.0, 0.0, 0.0, 0.0}, sLb{0.0, 0.0, 0.0, 0.0},
        sb{0.0, 0.0, 0.0, 0.0};
    if (ifc) {
        const Q Fb{lam * (Ub.r - UbL.r), lam * (Ub.p - UbL.p), lam * (Ub.un - UbL.un),
                   lam * (Ub.ut - UbL.ut)};
        IfState Wb, Wtb;
        dtb += flux_adj(W, Wt, dt, Fb, Wb, Wtb);
        grp_adj(fpL, fm, sL, s, G, Wb, Wtb, fpLb, fmb, sLb, sb);
    }
    // the line state again (a reload, so that it is not held across grp_adj)
    asm volatile("" ::: "memory");
    const Q q = load_line(src, P, m);
    const Q qL = shr(q), qR = shl(q);
    const Q fpb = shl(fpLb), sbR = shl(sLb);
    Q qb{fmb.r + fpb.r, fmb.p + fpb.p, fmb.un + fpb.un, fmb.ut + fpb.ut};
    Q st{0.0, 0.0, 0.0, 0.0};
    if (sig) {
        st = Q{sb.r + sbR.r + L.hh * (fpb.r - fmb.r), sb.p + sbR.p + L.hh * (fpb.p - fmb.p),
               sb.un + sbR.un + L.hh * (fpb.un - fmb.un), sb.ut + sbR.ut + L.hh * (fpb.ut - fmb.ut)};
    }
    Q pL, pC, pR;
    slope_adj(qL, q, qR, L, st, pL, pC, pR);
    const Q fromR = shl(pL), fromL = shr(pR);
    const Q xp{park[lane], park[kWave + lane], park[2 * kWave + lane], park[3 * kWave + lane]};
    const Q cb = pick(upd, cons_adj(q, xp, G), xp);
    const Q acc = mode == 1 ? load_line(d.lam_out, P, m) : Q{0.0, 0.0, 0.0, 0.0};
    qb = Q{qb.r + pC.r + fromR.r + fromL.r + cb.r + acc.r, qb.p + pC.p + fromR.p + fromL.p + cb.p + acc.p,
           qb.un + pC.un + fromR.un + fromL.un + cb.un + acc.un,
           qb.ut + pC.ut + fromR.ut + fromL.ut + cb.ut + acc.ut};

    double *out = mode == 0 ? (m.ydir ? d.lamA1 : d.lamB1) : d.lam_out;
    if (own) store_line(out, P, m, qb, false);
    const double part_sum = wave_sum(own ? dtb : 0.0);
    if (lane == 0) gp(d.dtpart)[part] = part_sum;
}

// Each wave walks a strided list of tasks.  Two waves per SIMD: nothing is
// loaded ahead, and only the GRP adjoint's own inputs and outputs stay in
// registers across it (the line state is loaded again after it), which
// keeps the kernel within 256 registers without spills; the second wave
// hides the latencies one wave alone would wait on.
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(2, 2))) adj_sweep_kernel(AdjDesc *__restrict__ descs,
                                                        const int4 *__restrict__ tasks, int ntasks,
                                                        int mode, int part_base) {
    __shared__ double park[4][4 * kWave];  // per wave: 4 doubles per lane
    const int lane = threadIdx.x & (kWave - 1);
    const int nwv = gridDim.x * (blockDim.x / kWave);
    for (int tk = __builtin_amdgcn_readfirstlane(xcd_block() * (blockDim.x / kWave) + threadIdx.x / kWave);
         tk < ntasks; tk += nwv) {
        const int4 t = uniform_task(tasks, tk);
        const LineMap m = line_map(descs[t.x], t, lane, kAdjHalo, kAdjOwn);
        sweep_task(descs[t.x], m, lane, mode, part_base + tk, park[threadIdx.x / kWave]);
    }
}

// ---------------------------------------------------------------- R4
// lam[argmax] += (sum of dt sensitivities) * d dt / d U[argmax], where
// dt = CFL h / s_max and s_max = speed(U[argmax]).  One block per instance;
// the sum runs in a fixed order (four strided partial sums per thread, a
// butterfly over each wave, then one wave over the wave sums), so the
// result is reproducible.  The arg-max cell's state and output adjoint are
// fetched by the first lane before the partials are summed.  A step whose
// arg-max was not found (no cell's speed equals the forward loop's s_max)
// adds nothing and sets status bit 1.
__global__ void __launch_bounds__(1024) adj_dt_term_kernel(AdjDesc *__restrict__ descs) {
    __shared__ double red[1024 / kWave];
    AdjDesc &d = descs[blockIdx.x];
    const int T = blockDim.x;
    const int lane = threadIdx.x & (kWave - 1), wid = threadIdx.x / kWave;
    const long long P = d.plane;
    long long o = -1;
    Q g{0.0, 0.0, 0.0, 0.0}, lam{0.0, 0.0, 0.0, 0.0};
    if (threadIdx.x == 0) {
        const long long k = d.ctl->argmax;
        if (k >= 0 && k < (long long)d.nx * d.ny) {
            o = (k / d.nx) * d.pitch + k % d.nx;
            g = speed_grad(d.U[o], d.U[P + o], d.U[2 * P + o], d.U[3 * P + o], d.gas);
            lam = Q{d.lam_out[o], d.lam_out[P + o], d.lam_out[2 * P + o], d.lam_out[3 * P + o]};
        }
    }
    double a0 = 0.0, a1 = 0.0, a2 = 0.0, a3 = 0.0;
    for (int r = 0; r < 3; r++) {
        const int lo = d.part_lo[r], hi = d.part_hi[r];
        int p = lo + (int)threadIdx.x;
        for (; p + 3 * T < hi; p += 4 * T) {
            a0 += d.dtpart[p];
            a1 += d.dtpart[p + T];
            a2 += d.dtpart[p + 2 * T];
            a3 += d.dtpart[p + 3 * T];
        }
        if (p < hi) a0 += d.dtpart[p];
        if (p + T < hi) a1 += d.dtpart[p + T];
        if (p + 2 * T < hi) a2 += d.dtpart[p + 2 * T];
    }
    double a = (a0 + a1) + (a2 + a3);
    for (int w = kWave / 2; w > 0; w >>= 1) a += __shfl_xor(a, w, kWave);  // every lane: the wave's sum
    if (lane == 0) red[wid] = a;
    __syncthreads();
    if (wid == 0) {
        const int nw = T / kWave;
        double b = lane < nw ? red[lane] : 0.0;
        for (int w = kWave / 2; w > 0; w >>= 1) b += __shfl_xor(b, w, kWave);
        if (lane == 0) {
            const double dtb = b;
            if (o >= 0) {
                const double f = dtb * (-d.fctl->dt / d.fctl->smax);
                d.lam_out[o] = lam.r + f * g.r;
                d.lam_out[P + o] = lam.p + f * g.p;
                d.lam_out[2 * P + o] = lam.un + f * g.un;
                d.lam_out[3 * P + o] = lam.ut + f * g.ut;
            } else {
                atomicOr(&d.ctl->status, 1);
            }
            d.ctl->dtbar = dtb;
            d.ctl->argmax = kNoCell;  // for the next reverse step's search
        }
    }
}

// ---------------------------------------------------------------- boundary conversions
// API planes are contiguous [4][ny][nx] in the order (rho, rho u, rho v, E);
// internal planes are pitched (rho, p, u, v).
__global__ void cons_to_prim_kernel(const double *__restrict__ cons, double *__restrict__ prim, int nx,
                                    int ny, int pitch, long long plane, double g1) {
    const long long n = (long long)nx * ny;
    const long long k = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= n) return;
    const int j = (int)(k / nx), i = (int)(k % nx);
    const long long o = (long long)j * pitch + i;
    const double r = cons[k], mu = cons[n + k], mv = cons[2 * n + k], E = cons[3 * n + k];
    const double u = mu / r, v = mv / r;
    prim[o] = r;
    prim[plane + o] = g1 * (E - 0.5 * (mu * u + mv * v));
    prim[2 * plane + o] = u;
    prim[3 * plane + o] = v;
}

__global__ void prim_to_cons_kernel(const double *__restrict__ prim, double *__restrict__ cons, int nx,
                                    int ny, int pitch, long long plane, double g1i) {
    const long long n = (long long)nx * ny;
    const long long k = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= n) return;
    const int j = (int)(k / nx), i = (int)(k % nx);
    const long long o = (long long)j * pitch + i;
    const double r = prim[o], p = prim[plane + o], u = prim[2 * plane + o], v = prim[3 * plane + o];
    cons[k] = r;
    cons[n + k] = r * u;
    cons[2 * n + k] = r * v;
    cons[3 * n + k] = p * g1i + 0.5 * r * (u * u + v * v);
}

// lam_prim = (d cons / d prim)^T g_cons at the primitive state prim.
__global__ void cons_grad_to_prim_kernel(const double *__restrict__ prim, const double *__restrict__ g,
                                         double *__restrict__ lam, int nx, int ny, int pitch,
                                         long long plane, double g1i) {
    const long long n = (long long)nx * ny;
    const long long k = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= n) return;
    const int j = (int)(k / nx), i = (int)(k % nx);
    const long long o = (long long)j * pitch + i;
    const double r = prim[o], u = prim[2 * plane + o], v = prim[3 * plane + o];
    const double gr = g[k], gmu = g[n + k], gmv = g[2 * n + k], gE = g[3 * n + k];
    lam[o] = gr + u * gmu + v * gmv + 0.5 * (u * u + v * v) * gE;
    lam[plane + o] = g1i * gE;
    lam[2 * plane + o] = r * gmu + r * u * gE;
    lam[3 * plane + o] = r * gmv + r * v * gE;
}

// g_cons = (d prim / d cons)^T lam_prim at the conserved input state.
__global__ void prim_grad_to_cons_kernel(const double *__restrict__ cons, const double *__restrict__ lam,
                                         double *__restrict__ g, int nx, int ny, int pitch,
                                         long long plane, double g1) {
    const long long n = (long long)nx * ny;
    const long long k = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= n) return;
    const int j = (int)(k / nx), i = (int)(k % nx);
    const long long o = (long long)j * pitch + i;
    const double r = cons[k], u = cons[n + k] / r, v = cons[2 * n + k] / r;
    const double lr = lam[o], lp = lam[plane + o], lu = lam[2 * plane + o], lv = lam[3 * plane + o];
    g[k] = lr - (lu * u + lv * v) / r + lp * g1 * 0.5 * (u * u + v * v);
    g[n + k] = lu / r - lp * g1 * u;
    g[2 * n + k] = lv / r - lp * g1 * v;
    g[3 * n + k] = lp * g1;
}

// ---------------------------------------------------------------- launchers
static inline unsigned blocks_for(long long n, int per) { return (unsigned)((n + per - 1) / per); }

void adj_launch_fwd(AdjDesc *descs, const int4 *tasks, int ntasks, int max_waves, hipStream_t s) {
    const int waves = ntasks < max_waves ? ntasks : max_waves;
    hipLaunchKernelGGL(adj_fwd_kernel, dim3(blocks_for(waves, 4)), dim3(256), 0, s, descs, tasks, ntasks);
}
void adj_launch_fwd_y(AdjDesc *descs, const int4 *tasks, int ntasks, hipStream_t s) {
    if (ntasks > 0)
        hipLaunchKernelGGL(adj_fwd_y_kernel, dim3(blocks_for(ntasks, 4)), dim3(256), 0, s, descs,
                           tasks, ntasks);
}
__global__ void __launch_bounds__(256) adj_copy_kernel(const double *const *__restrict__ src,
                                                       double *const *__restrict__ dst,
                                                       const long long *__restrict__ count) {
    const int i = blockIdx.y;
    const double2 *__restrict__ a = reinterpret_cast<const double2 *>(src[i]);
    double2 *__restrict__ b = reinterpret_cast<double2 *>(dst[i]);
    const long long n2 = count[i] / 2;
    const long long stride = (long long)gridDim.x * blockDim.x;
    long long k = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    // four 16-byte loads in flight per lane before their stores
    for (; k + 3 * stride < n2; k += 4 * stride) {
        const double2 x0 = a[k], x1 = a[k + stride], x2 = a[k + 2 * stride], x3 = a[k + 3 * stride];
        b[k] = x0;
        b[k + stride] = x1;
        b[k + 2 * stride] = x2;
        b[k + 3 * stride] = x3;
    }
    for (; k < n2; k += stride) b[k] = a[k];
}

void adj_launch_copy(const double *const *src, double *const *dst, const long long *count, int n,
                     long long max_count, hipStream_t s) {
    // about 4096 blocks over all meshes, each lane moving at least four 16-byte pieces
    const long long per_mesh = std::max(16, 4096 / std::max(1, n));
    const unsigned bx = (unsigned)std::max<long long>(1, std::min<long long>(per_mesh, (max_count / 2 + 1023) / 1024));
    hipLaunchKernelGGL(adj_copy_kernel, dim3(bx, n), dim3(256), 0, s, src, dst, count);
}
void adj_launch_sweep(AdjDesc *descs, const int4 *tasks, int ntasks, int mode, int part_base,
                      int max_waves, hipStream_t s) {
    // max_waves: the resident waves wanted (two per SIMD fit)
    const int waves = ntasks < max_waves ? ntasks : max_waves;
    hipLaunchKernelGGL(adj_sweep_kernel, dim3(blocks_for(waves, 4)), dim3(256), 0, s, descs, tasks,
                       ntasks, mode, part_base);
}
void adj_launch_dt_term(AdjDesc *descs, int ninst, int nparts, hipStream_t s) {
    (void)nparts;
    hipLaunchKernelGGL(adj_dt_term_kernel, dim3(ninst), dim3(1024), 0, s, descs);
}
void adj_launch_cons_to_prim(const double *cons, double *prim, int nx, int ny, int pitch, long long plane,
                             double g1, hipStream_t s) {
    hipLaunchKernelGGL(cons_to_prim_kernel, dim3(blocks_for((long long)nx * ny, 256)), dim3(256), 0, s,
                       cons, prim, nx, ny, pitch, plane, g1);
}
void adj_launch_prim_to_cons(const double *prim, double *cons, int nx, int ny, int pitch, long long plane,
                             double g1i, hipStream_t s) {
    hipLaunchKernelGGL(prim_to_cons_kernel, dim3(blocks_for((long long)nx * ny, 256)), dim3(256), 0, s,
                       prim, cons, nx, ny, pitch, plane, g1i);
}
void adj_launch_cons_grad_to_prim(const double *prim, const double *g, double *lam, int nx, int ny,
                                  int pitch, long long plane, double g1i, hipStream_t s) {
    hipLaunchKernelGGL(cons_grad_to_prim_kernel, dim3(blocks_for((long long)nx * ny, 256)), dim3(256), 0,
                       s, prim, g, lam, nx, ny, pitch, plane, g1i);
}
void adj_launch_prim_grad_to_cons(const double *cons, const double *lam, double *g, int nx, int ny,
                                  int pitch, long long plane, double g1, hipStream_t s) {
    hipLaunchKernelGGL(prim_grad_to_cons_kernel, dim3(blocks_for((long long)nx * ny, 256)), dim3(256), 0,
                       s, cons, lam, g, nx, ny, pitch, plane, g1);
}

}  // namespace fvm2d
