// Descriptors and launchers of the reverse-step kernels (fvm2d_adjoint.hip).
#pragma once

#include <hip/hip_runtime.h>

#include "fvm2d_device.h"

namespace fvm2d {

// Per-instance scalars of one reverse step, resident on the device.
struct AdjCtl {
    double dtbar;      // total sensitivity of the loss to dt_n
    long long argmax;  // linear index (j * nx + i) of the first cell attaining s_max
                       // (kNoCell between reverse steps)
    int status;        // bit 0: non-finite wave speed, or no cell attaining s_max
    int pad;
};
constexpr long long kNoCell = 0x7fffffffffffffffll;

// Task of the line kernels: x = instance, y = 0 for an X line (a row) or 1
// for a Y line (a column), z = window along the line, w = row or column.
struct AdjDesc {
    int nx, ny, pitch, strategy;
    long long plane;
    double cfl, hmin;
    Line lx, ly;
    Gas gas;
    const double *U;       // state U_n (pitched primitive planes)
    const Ctl *fctl;       // the forward time loop's control block of U_n: its
                           // s_max and dt_n = CFL h / s_max (no t_end clamp)
    double *A1, *B1;       // first sweeps of the two branches, recomputed
    const double *lam_in;  // dL/dU_{n+1}
    double *lamA1, *lamB1; // dL/dA1, dL/dB1
    double *lam_out;       // dL/dU_n
    double *dtpart;        // per-task dt sensitivities (shared array)
    AdjCtl *ctl;
    int part_lo[3], part_hi[3];    // this instance's ranges in dtpart
};

// max_waves: the waves of a launch (each walks a strided task list).
void adj_launch_fwd(AdjDesc *descs, const int4 *tasks, int ntasks, int max_waves, hipStream_t s);
// B1 = Y(U) by the register march; tasks {instance, 64-column window, r0, r1}
// tile every instance's interior rows [2, ny-2) for every window.
void adj_launch_fwd_y(AdjDesc *descs, const int4 *tasks, int ntasks, hipStream_t s);
// Copy the states of all n instances from one location to another (one
// launch): src[i] -> dst[i], count[i] doubles each (even, 16-byte aligned).
void adj_launch_copy(const double *const *src, double *const *dst, const long long *count, int n,
                     long long max_count, hipStream_t s);
void adj_launch_sweep(AdjDesc *descs, const int4 *tasks, int ntasks, int mode, int part_base,
                      int max_waves, hipStream_t s);
void adj_launch_dt_term(AdjDesc *descs, int ninst, int nparts, hipStream_t s);
void adj_launch_cons_to_prim(const double *cons, double *prim, int nx, int ny, int pitch,
                             long long plane, double g1, hipStream_t s);
void adj_launch_prim_to_cons(const double *prim, double *cons, int nx, int ny, int pitch,
                             long long plane, double g1i, hipStream_t s);
void adj_launch_cons_grad_to_prim(const double *prim, const double *g, double *lam, int nx, int ny,
                                  int pitch, long long plane, double g1i, hipStream_t s);
void adj_launch_prim_grad_to_cons(const double *cons, const double *lam, double *g, int nx, int ny,
                                  int pitch, long long plane, double g1, hipStream_t s);

}  // namespace fvm2d
