// Wave-level building blocks shared by the forward and adjoint kernels:
// DPP lane shifts across the 64-wide wavefront and the lane-parallel
// X-sweep (lanes = consecutive cells of one line).
#pragma once

#include <hip/hip_runtime.h>

#include "fvm2d_cell.h"

namespace fvm2d {

__device__ __forceinline__ double dpp_shr1(double x) {  // lane l <- lane l-1
    int lo = __double2loint(x), hi = __double2hiint(x);
    lo = __builtin_amdgcn_update_dpp(lo, lo, 0x138, 0xf, 0xf, false);
    hi = __builtin_amdgcn_update_dpp(hi, hi, 0x138, 0xf, 0xf, false);
    return __hiloint2double(hi, lo);
}
__device__ __forceinline__ double dpp_shl1(double x) {  // lane l <- lane l+1
    int lo = __double2loint(x), hi = __double2hiint(x);
    lo = __builtin_amdgcn_update_dpp(lo, lo, 0x130, 0xf, 0xf, false);
    hi = __builtin_amdgcn_update_dpp(hi, hi, 0x130, 0xf, 0xf, false);
    return __hiloint2double(hi, lo);
}
__device__ __forceinline__ Q shr(const Q &q) {
    return Q{dpp_shr1(q.r), dpp_shr1(q.p), dpp_shr1(q.un), dpp_shr1(q.ut)};
}
__device__ __forceinline__ Q shl(const Q &q) {
    return Q{dpp_shl1(q.r), dpp_shl1(q.p), dpp_shl1(q.un), dpp_shl1(q.ut)};
}

// X-order (rho, p, u, v) <-> Y-line order (rho, p, v, u)
__device__ __forceinline__ Q swap_uv(const Q &q) { return Q{q.r, q.p, q.ut, q.un}; }

// ---------------------------------------------------------------- sweeps
// One X-sweep of a row held across the wave, lanes = consecutive columns.
// Output is valid on lanes 2..61 (a two-lane stencil halo each side).
template <bool Exact>
__device__ __forceinline__ Q xsweep(const Q &q, const Line &L, const Gas &G, double dt,
                                    double lam) {
    const Q qL = shr(q), qR = shl(q);
    const Q s = sweep_slope<Exact>(qL, q, qR, L);
    const Q fm = sweep_face_minus<Exact>(q, s, L);
    const Q sL = shr(s);
    // east face of the west neighbour, from its state (already shifted in)
    // and its shifted slope: the same operations as on that lane
    const Q fpL = sweep_face_plus<Exact>(qL, sL, L);
    IfState W, Wt;
    grp<Exact>(fpL, fm, sL, s, G, W, Wt);  // interface on the west face
    const Q F = flux<Exact>(W, Wt, sweep_k<Exact>(dt, L));
    const Q FR = shl(F);                   // interface on the east face
    return update<Exact>(q, F, FR, lam, G);
}

// Row exchange through a wave-private LDS buffer of 4 x 64 doubles: every
// lane stores its value in its own slot and reads its west / east
// neighbour's slot, clamped at the wave's edges (edge lanes read their own
// value, as the DPP shifts leave them).  The step kernel exchanges a row
// well before it needs the neighbours, so the LDS latency is hidden and
// the two VALU instructions per 32-bit half of a DPP shift are saved.
struct Xch {
    double *b;     // wave-private buffer
    int me, w, e;  // own, west and east slot
};
__device__ __forceinline__ Xch make_xch(double *b, int lane) {
    return Xch{b, lane, lane > 0 ? lane - 1 : 0, lane < kWave - 1 ? lane + 1 : kWave - 1};
}
__device__ __forceinline__ void xput(const Xch &x, const Q &q) {
    x.b[x.me] = q.r;
    x.b[kWave + x.me] = q.p;
    x.b[2 * kWave + x.me] = q.un;
    x.b[3 * kWave + x.me] = q.ut;
    __builtin_amdgcn_wave_barrier();
}
__device__ __forceinline__ Q xget(const Xch &x, int i) {
    return Q{x.b[i], x.b[kWave + i], x.b[2 * kWave + i], x.b[3 * kWave + i]};
}

// xsweep with the row's west / east neighbours already exchanged (qL, qR);
// the slope and flux exchanges stay DPP shifts.  Same operations and
// results as xsweep.
template <bool Exact>
__device__ __forceinline__ Q xsweep_pre(const Q &q, const Q &qL, const Q &qR, const Line &L,
                                        const Gas &G, double dt, double lam) {
    const Q s = sweep_slope<Exact>(qL, q, qR, L);
    const Q fm = sweep_face_minus<Exact>(q, s, L);
    const Q sL = shr(s);
    const Q fpL = sweep_face_plus<Exact>(qL, sL, L);
    IfState W, Wt;
    grp<Exact>(fpL, fm, sL, s, G, W, Wt);
    const Q F = flux<Exact>(W, Wt, sweep_k<Exact>(dt, L));
    const Q FR = shl(F);
    return update<Exact>(q, F, FR, lam, G);
}

}  // namespace fvm2d
