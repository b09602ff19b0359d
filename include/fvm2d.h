/*
 * fvm2d: MI355X (gfx950) 2D relaxation-GRP finite-volume Euler solver.
 * C ABI of libfvm2d.so.  Plain pointers and sizes only.
 *
 * State layout at this boundary is the reference's: an array of vec5d
 * records, [ny][nx][5] doubles, row 0 = north, column 0 = west, primitive
 * (rho, p, u, v, unused); the outer two rows/columns are ghost cells
 * (src/ghostCellPresets.cuh:4-45, src/wrappers.cuh:67-78, README.md:4).
 *
 * Every function returns 0 on success or a negative FVM2D_ERR_* code; the
 * message of the last failure on the calling thread is fvm2d_last_error().
 * The reference checks no return codes; here bad arguments, HIP failures
 * and a non-finite wave speed (NaN state) are reported.
 */
#ifndef FVM2D_H
#define FVM2D_H

#ifdef __cplusplus
extern "C" {
#endif

#define FVM2D_OK 0
#define FVM2D_ERR_ARG (-1)
#define FVM2D_ERR_HIP (-2)
#define FVM2D_ERR_NONFINITE (-3)
#define FVM2D_ERR_NODEVICE (-4)
#define FVM2D_ERR_STATE (-5)

/* enum class ghostCellStrategy, src/wrappers.cuh:17-24 (same order).
 * REFLECTIVE runs the double-Mach preset, as src/wrappers.cuh:277-280 does. */
enum {
    FVM2D_OUTFLOW = 0,
    FVM2D_REFLECTIVE = 1,
    FVM2D_JET = 2,
    FVM2D_JET800 = 3,
    FVM2D_DOUBLE_MACH = 4
};

/* Arithmetic mode. FAST shares reciprocals and uses FMA; EXACT evaluates the
 * reference's expressions in its order without contraction and reproduces
 * the FP64 host restatement bit for bit. */
enum { FVM2D_MODE_FAST = 0, FVM2D_MODE_EXACT = 1 };

/* The solver's settable parameters: RGRP_2D_CUDA's setters
 * (src/wrappers.cuh:110-152) and defaults (src/wrappers.cuh:49-52). */
typedef struct {
    double width;   /* setSpatialLayout, default 1 */
    double height;  /* setSpatialLayout, default 1 */
    double t_end;   /* setTimeLayout, default 0 */
    double cfl;     /* setCFL / setTimeLayout, default 0.5 */
    double alpha;   /* setAlpha, default 1.9 */
    double gamma;   /* setGamma, default 1.4 */
    int strategy;   /* setGhostCellStrategy, default FVM2D_OUTFLOW */
    int max_steps;  /* extension: < 0 runs until t >= t_end (reference behaviour) */
    int mode;       /* FVM2D_MODE_FAST (default) or FVM2D_MODE_EXACT */
    int reserved;
} fvm2d_params;

/* Fill p with the reference's defaults (src/wrappers.cuh:49-52). */
void fvm2d_params_default(fvm2d_params *p);

/* Message of the last failing call on this thread ("" if none). */
const char *fvm2d_last_error(void);

/* 1 if a HIP device is usable, else 0. */
int fvm2d_device_available(void);

/* RGRP_2D_CUDA::solve (src/wrappers.cuh:154-306): blocking; advances U5 in
 * place from t = 0 to t_end (or max_steps), leaves the final state with its
 * ghost cells filled.  steps_out: time steps taken; ms_out: milliseconds of
 * the device time loop (the region of src/wrappers.cuh:193-288). */
int fvm2d_solve(double *U5, int nx, int ny, const fvm2d_params *p, int *steps_out,
                double *ms_out);

/* No counterpart in the reference; for callers that cannot allocate device
 * memory (stream-capture set-up, serving loops, harnesses that forbid
 * hipMalloc).  The solve loop of fvm2d_solve run in place on the caller's
 * device planes [4][ny][nx] (rho, p, u, v, contiguous, row 0 = north,
 * ghosts included), with every other buffer in a library-owned device array
 * of fixed size (33 MiB per GPU, held until the process ends): no
 * hipMalloc, hipFree or other allocating call.  Meshes up to 1048576 cells
 * (larger: FVM2D_ERR_ARG).  Blocking, on the null stream.  All calls on one
 * GPU share that array: calls that overlap on one GPU race. */
int fvm2d_solve_device_noalloc(double *d_planes, int nx, int ny, const fvm2d_params *p,
                               int *steps_out, double *t_out);

/* ---- resident solver: the state stays in HBM between calls ------------- */
typedef struct fvm2d_solver fvm2d_solver;

/* RGRP_2D_CUDA(U0) (src/wrappers.cuh:67-78) split in create + upload. */
int fvm2d_create(int nx, int ny, const fvm2d_params *p, fvm2d_solver **out);
int fvm2d_set_params(fvm2d_solver *s, const fvm2d_params *p);
int fvm2d_upload(fvm2d_solver *s, const double *U5);
/* Device-to-device copy in/out of a contiguous [4][ny][nx] plane array
 * (rho, p, u, v) that lives in device memory. */
int fvm2d_upload_planes_device(fvm2d_solver *s, const double *d_planes);
int fvm2d_download_planes_device(fvm2d_solver *s, double *d_planes);
/* The solve loop of src/wrappers.cuh:191-288 on the resident state: t
 * restarts at 0, the state continues from the last one (T11 of the
 * reference: a second solve() resumes from the previous result). */
int fvm2d_run(fvm2d_solver *s, int *steps_out, double *t_out, double *ms_out);
/* solve() copies the final state out: src/wrappers.cuh:290-291. */
int fvm2d_download(fvm2d_solver *s, double *U5);
/* Record mode (src/wrappers.cuh:253-261): inside a time loop started by
 * fvm2d_begin, after at least one step, the state after the last completed
 * step as the reference copies it out for a frame -- after the branch
 * average, before that step's ghost fill (the fill runs after the copy,
 * src/wrappers.cuh:265-281).  Layout as fvm2d_download.  FVM2D_ERR_STATE
 * outside a time loop or before its first step. */
int fvm2d_download_frame(fvm2d_solver *s, double *U5);
/* No counterpart in the reference (its frame writer does this on the host
 * with OpenCV, src/wrappers.cuh:323-340): the density of the frame of
 * fvm2d_download_frame as the 8-bit image the reference's frame file holds
 * before its colour map, computed on the device so that nx * ny bytes cross
 * instead of the state.  gray: ny x nx bytes, row 0 first.  lo and hi are
 * the minimum and maximum density over every cell (ghosts included), s =
 * 255 / (hi - lo), pixel = rint(rho s - lo s) clamped to 0..255 (round half
 * to even, each operation rounded once); all pixels 0 when hi == lo; NaN
 * densities are left out of lo and hi and give 0.  Errors as
 * fvm2d_download_frame. */
int fvm2d_download_frame_gray(fvm2d_solver *s, unsigned char *gray);
/* Total time steps over all runs (getTimeStagesCount, src/wrappers.cuh:313-316). */
int fvm2d_steps_total(const fvm2d_solver *s);
int fvm2d_destroy(fvm2d_solver *s);

/* Launch geometry of the step kernel.  strip_rows: output rows per wave
 * (0 keeps the current value); waves_per_block: 1..4 (0 keeps);
 * persistent: 1 = persistent grid with dynamic task tickets, 0 = one wave
 * per task, negative keeps the current choice.  Results do not depend on
 * the geometry. */
int fvm2d_set_launch(fvm2d_solver *s, int strip_rows, int waves_per_block, int persistent);
int fvm2d_get_launch(const fvm2d_solver *s, int *strip_rows, int *waves_per_block,
                     int *persistent);
/* Time every candidate launch geometry on a scratch copy of the state and
 * keep the fastest; the state is left unchanged.  best_ms_out: per step. */
int fvm2d_autotune(fvm2d_solver *s, int steps_per_trial, double *best_ms_out);
/* Start a time loop at t = 0 without running it, then advance at most
 * nsteps steps (asynchronously); fvm2d_sync waits and reports. */
int fvm2d_begin(fvm2d_solver *s);
int fvm2d_advance(fvm2d_solver *s, int nsteps);
int fvm2d_sync(fvm2d_solver *s, int *steps_out, double *t_out, double *dt_out);
/* Wall time of nsteps launches bracketed by HIP events on the solver's
 * stream (after fvm2d_begin).  ms_out: total milliseconds. */
int fvm2d_time_steps(fvm2d_solver *s, int nsteps, double *ms_out);
/* Step-kernel launch duration on the solver's stream: avg = elapsed time of
 * nsteps back-to-back launches between two HIP events / nsteps; min = the
 * shortest of nsteps further launches each bracketed by its own events.
 * Both in ms; 2 nsteps steps are advanced. */
int fvm2d_time_kernel(fvm2d_solver *s, int nsteps, double *avg_ms_out, double *min_ms_out);

/* ---- batched forward (no counterpart in the reference) ----------------- *
 * n independent meshes of their own size and parameters advance together:
 * one launch per time step covers every instance's tiles (persistent waves
 * taking task tickets, or one wave per task: the autotuner chooses), each
 * instance keeps its own t, dt and step counter on the device, and the
 * step loop is captured in a hipGraph. */
typedef struct fvm2d_batch fvm2d_batch;
int fvm2d_batch_create(int n, const int *nx, const int *ny, const fvm2d_params *p,
                       fvm2d_batch **out);
int fvm2d_batch_upload(fvm2d_batch *b, int i, const double *U5);
int fvm2d_batch_begin(fvm2d_batch *b);
/* Launch up to nsteps steps (captured once per chunk size in a hipGraph);
 * returns when every instance is done or nsteps launches were issued. */
int fvm2d_batch_run(fvm2d_batch *b, int nsteps, double *ms_out);
int fvm2d_batch_status(fvm2d_batch *b, int i, int *steps_out, double *t_out);
int fvm2d_batch_download(fvm2d_batch *b, int i, double *U5);
int fvm2d_batch_set_launch(fvm2d_batch *b, int strip_rows, int waves_per_block, int persistent);
int fvm2d_batch_autotune(fvm2d_batch *b, int steps_per_trial, double *best_ms_out);
int fvm2d_batch_destroy(fvm2d_batch *b);

/* Diagnostic with no counterpart in the reference: the strips a launch of
 * one nx x ny mesh assigns to its waves, in launch position order, as
 * (column window, first row, end row) triples in out[3 k .. 3 k + 2] for
 * k < min(*ntasks_out, max_tasks).  cu_count and blocks_per_cu (resident
 * blocks of waves_per_block waves per CU) describe the device; no device
 * is used. */
int fvm2d_strip_layout(int nx, int ny, int strip_rows, int waves_per_block, int persistent,
                       int cu_count, int blocks_per_cu, int *out, int max_tasks, int *ntasks_out);

/* ---- adjoint (no counterpart in the reference) ------------------------- *
 * Hand-written reverse mode of the full split step (limiter and GRP
 * differentiated along the branch taken, dt differentiated through its
 * arg-max cell) over nsteps fixed steps dt = CFL h / s_max (t_end is not
 * used), with binomial time-step checkpointing in checkpoint_slots states
 * per instance (slot 0 holds the initial state; <= 0 stores every step).
 * States and gradients cross the boundary as device pointers to contiguous
 * [4][ny][nx] conserved planes in the order (rho, rho u, rho v, E); one
 * pointer per instance.  Arithmetic is the FAST mode. */
typedef struct fvm2d_adj fvm2d_adj;
typedef struct {
    double forward_ms;       /* the forward call (with its checkpoint copies) */
    double backward_ms;      /* the backward call (reverse steps + recomputation) */
    double forward_step_ms;  /* forward_ms / nsteps */
    double reverse_step_ms;  /* mean duration of one reverse step */
    int reverse_steps;
    int recompute_steps;     /* forward steps re-run by the backward call */
    int checkpoint_slots;
    int schedule_cost;       /* forward steps of the schedule after the stored state */
} fvm2d_adj_stats;
int fvm2d_adj_create(int n, const int *nx, const int *ny, const fvm2d_params *p, int nsteps,
                     int checkpoint_slots, fvm2d_adj **out);
int fvm2d_adj_forward(fvm2d_adj *a, const double *const *d_U0_cons, double *const *d_UN_cons);
int fvm2d_adj_backward(fvm2d_adj *a, const double *const *d_gN_cons, double *const *d_g0_cons);
/* Same with host pointers: the library stages the planes through its own
 * device buffers (the timings in fvm2d_adj_stats exclude the copies). */
int fvm2d_adj_forward_host(fvm2d_adj *a, const double *const *U0_cons, double *const *UN_cons);
int fvm2d_adj_backward_host(fvm2d_adj *a, const double *const *gN_cons, double *const *g0_cons);
int fvm2d_adj_get_stats(const fvm2d_adj *a, fvm2d_adj_stats *out);
/* Launch geometry of the forward step kernel (as fvm2d_set_launch).  Without
 * it the first forward call autotunes the geometry on its initial state. */
int fvm2d_adj_set_launch(fvm2d_adj *a, int strip_rows, int waves_per_block, int persistent);
int fvm2d_adj_destroy(fvm2d_adj *a);

#ifdef __cplusplus
}
#endif
#endif /* FVM2D_H */
