// Register march of a Y-sweep down a strip of rows (lanes = columns), shared
// by the step kernel and the adjoint's forward recompute.
#pragma once

#include "fvm2d_cell.h"

namespace fvm2d {

// Register pipeline of a Y-sweep marching south (increasing row j).  In the
// sweep's line order the "left" neighbour of row j is row j+1, which is the
// orientation the reference's rotate_and_flip_vertically produces.
struct YMarch {
    Q qj, qj1;   // rows j and j+1 (line order)
    Q sj;        // slope of row j (its south face is recomputed from it)
    Q Fprev;     // flux through the north face of row j
};

template <bool Exact>
__device__ __forceinline__ void ymarch_init(YMarch &m, const Q &m2, const Q &m1, const Q &q0,
                                            const Q &p1, const Line &L, const Gas &G, double dt) {
    // rows r0-2, r0-1, r0, r0+1
    const Q s_m1 = sweep_slope<Exact>(q0, m1, m2, L);
    const Q s_0 = sweep_slope<Exact>(p1, q0, m1, L);
    const Q fm_m1 = sweep_face_minus<Exact>(m1, s_m1, L);
    const Q fp_0 = sweep_face_plus<Exact>(q0, s_0, L);
    IfState W, Wt;
    grp<Exact>(fp_0, fm_m1, s_0, s_m1, G, W, Wt);
    m.Fprev = flux<Exact>(W, Wt, sweep_k<Exact>(dt, L));
    m.qj = q0;
    m.qj1 = p1;
    m.sj = s_0;
}

// Consume row j+2, return the updated row j.
template <bool Exact>
__device__ __forceinline__ Q ymarch_step(YMarch &m, const Q &qn, const Line &L, const Gas &G,
                                         double dt, double lam) {
    const Q s1 = sweep_slope<Exact>(qn, m.qj1, m.qj, L);
    const Q fp1 = sweep_face_plus<Exact>(m.qj1, s1, L);
    const Q fmj = sweep_face_minus<Exact>(m.qj, m.sj, L);
    IfState W, Wt;
    grp<Exact>(fp1, fmj, s1, m.sj, G, W, Wt);
    const Q F = flux<Exact>(W, Wt, sweep_k<Exact>(dt, L));  // south face of row j
    const Q out = update<Exact>(m.qj, F, m.Fprev, lam, G);
    m.qj = m.qj1;
    m.qj1 = qn;
    m.sj = s1;
    m.Fprev = F;
    return out;
}

}  // namespace fvm2d
