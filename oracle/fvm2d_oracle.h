/*
 * FP64 host oracle for the 2D relaxation-GRP finite-volume Euler solver.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker: only tests/,
 * __graft_entry__.smoke() and bench.py's cpu_baseline leg load it.  The
 * product library (libfvm2d.so) never links, loads or calls it.
 *
 * It restates, statement for statement and in the same floating-point
 * operation order, the reference's numerical path:
 *   src/ghostCellPresets.cuh:4-45    vec5d (4-slot arithmetic, slot 4 -> 0)
 *   src/ghostCellPresets.cuh:47-241  the four ghost-cell presets
 *   src/basicFuns.cuh:14-34          flux
 *   src/basicFuns.cuh:37-72          MUSCL_rel_RP_GRP (limited slopes + faces)
 *   src/basicFuns.cuh:75-192         rel_RP_posi_fix (relaxation GRP)
 *   src/basicFuns.cuh:195-212        primitive <-> conserved
 *   src/basicFuns.cuh:215-226        maxPropagtingSpeed
 *   src/basicFuns.cuh:229-246        forward_x_dir
 *   src/basicFuns.cuh:249-262        rotate_and_flip_vertically
 *   src/basicFuns.cuh:264-274        combineResults
 *   src/wrappers.cuh:154-306         RGRP_2D_CUDA::solve time loop
 * Built with g++ -O2 -ffp-contract=off (no FMA contraction) so every
 * operation is a single IEEE-754 double rounding.
 *
 * Layout at this boundary: the reference's vec5d array, [ny][nx][5] doubles,
 * row 0 = north, primitive (rho, p, u, v, unused).
 */
#ifndef FVM2D_ORACLE_H
#define FVM2D_ORACLE_H

#ifdef __cplusplus
extern "C" {
#endif

/* Same field order and meaning as fvm2d_params in include/fvm2d.h. */
typedef struct {
    double width, height, t_end, cfl, alpha, gamma;
    int strategy;   /* 0 outflow, 1 reflective(=double Mach), 2 jet, 3 jet800, 4 double Mach */
    int max_steps;  /* < 0: run until t >= t_end (reference behaviour) */
    int mode;       /* ignored by the oracle */
    int reserved;
} orc_params;

/* Full solve, in place on U5 ([ny][nx][5]).  Returns 0 on success. */
int orc_solve(double *U5, int nx, int ny, const orc_params *p, int nthreads,
              int *steps_out, double *t_out, double *ms_out);

/* Record mode (wrappers.cuh:253-261, setRecordModel(true, gap)): the same
 * solve, and the state as the reference copies it for a frame -- after the
 * branch average, before the ghost fill -- at the first step and every gap
 * steps after it: frames5 receives up to max_frames [ny][nx][5] states,
 * frame_index their file indices; *nframes_out counts every frame taken. */
int orc_solve_record(double *U5, int nx, int ny, const orc_params *p, int nthreads, int gap,
                     double *frames5, int max_frames, int *nframes_out, int *frame_index,
                     int *steps_out, double *t_out);

/* Per-routine entry points for unit tests (line order: left, centre, right). */
void orc_slope(const double *qL, const double *q, const double *qR, double h,
               double alpha, double *sigma4);
void orc_grp(const double *faceL4, const double *faceR4, const double *slopeL4,
             const double *slopeR4, double gamma, double *W5, double *Wt5, int *branch);
void orc_flux(const double *W5, const double *Wt5, double k, double *F4);
void orc_cons(const double *prim4, double gamma, double *cons4);
void orc_prim(const double *cons4, double gamma, double *prim4);
double orc_speed(const double *prim4, double gamma);
/* Two-phase ghost fill of a [ny][nx][5] array at time t. */
void orc_fill(double *U5, int nx, int ny, int strategy, double width,
              double height, double t);
/* One directional sweep (X on the given array), in place. */
void orc_sweep_x(double *U5, int nx, int ny, double width, double dt,
                 double alpha, double gamma, int nthreads);
/* Anti-transpose with u<->v swap: out is [nx][ny][5]. */
void orc_rotate(const double *in5, double *out5, int nx, int ny);

/* Host reference of the differentiated solve (fvm2d_oracle_adjoint.cpp):
 * nsteps steps with dt = CFL h / s_max, FAST per-cell numerics; states and
 * gradients are conserved planes [4][ny][nx] (rho, rho u, rho v, E). */
int orc_adj_forward(const double *U0cons, int nx, int ny, const orc_params *p, int nsteps,
                    double *UNcons);
int orc_adj_backward(const double *U0cons, int nx, int ny, const orc_params *p, int nsteps,
                     const double *gNcons, double *g0cons);

#ifdef __cplusplus
}
#endif
#endif
