// Host-side launchers of the step kernels (defined in fvm2d_step.hip).
#pragma once

#include <hip/hip_runtime.h>

#include "fvm2d_device.h"

namespace fvm2d {

// One time step of ninst instances: the step kernel (a task is one wave's
// strip: {instance, column window, first row, end row}), then the advance
// kernel that moves each instance's control block to the next step.
void launch_step(bool exact, const InstDesc *descs, Ctl *ctls, int ninst, const int4 *tasks, int ntasks,
                 Sched *sched, int persistent, int grid, int wpb, hipStream_t stream);
void launch_speed(bool exact, const InstDesc *descs, Ctl *ctls, int inst, int use_buf1,
                  long long nrows, hipStream_t stream);
void launch_ctl_start(const InstDesc *descs, Ctl *ctls, int ninst, const double *t0s,
                      hipStream_t stream);
void launch_ctl_restore(Ctl *ctls, const Ctl *saved, int ninst, hipStream_t stream);
// Record-mode frame of one mesh (the state before the last step's ghost
// fill) from that step's input prev and output cur, into out (4 planes).
void launch_frame(bool exact, const InstDesc &d, const double *prev, const double *cur, double dt,
                  double *out, hipStream_t stream);
// The 8-bit image of a frame's density (min-max normalised, rounded):
// gray[j * nx + i]; range holds two 64-bit words of scratch.
void launch_frame_gray(const InstDesc &d, const double *frame, unsigned long long *range,
                       unsigned char *gray, hipStream_t stream);
int step_kernel_occupancy(bool exact, int wpb);

}  // namespace fvm2d
