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

}  // namespace fvm2d
