// Device-side descriptors shared by the step kernels and the host runtime.
#pragma once

#include <cstdint>

#include "fvm2d_cell.h"

namespace fvm2d {

constexpr int kWave = 64;          // wavefront width on gfx950
constexpr int kOwn = kWave - 4;    // output columns per wave window (2-cell halo each side)

enum Strategy : int {
    kOutflow = 0,
    kReflective = 1,  // runs the double-Mach preset, as the reference does
    kJet = 2,
    kJet800 = 3,
    kDoubleMach = 4,
};

// One mesh ("instance").  State lives in two ping-pong buffers of four
// planes each (rho, p, u, v), row-major with a padded row pitch; the buffer
// holding the current state is buf[steps & 1].
struct InstDesc {
    double *buf0;
    double *buf1;
    long long plane;   // doubles per plane = pitch * ny
    int nx, ny, pitch;
    int nw;            // column windows: ceil((nx - 4) / kOwn)
    int ntask;         // tasks of this instance in the launch layout
    int R;             // target strip height (output rows per wave)
    int strategy;
    int max_steps;     // < 0: no step limit
    int task_base;     // first task index of this instance
    int pad0;
    double t_end, cfl, hmin;  // hmin = min(width/nx, height/ny)
    Line lx, ly;
    Gas gas;
};

// Per-instance time-loop control, resident on the device.  A step launch
// only reads it (and folds its waves' speed maxima into smax_next); the
// advance kernel that follows every step launch moves it to the next step.
struct Ctl {
    double t;                      // current time
    double dt;                     // step size for the next step
    double dt_used;                // step size of the last step taken
    double smax;                   // max |(u,v)| + c over the current state
    unsigned long long smax_next;  // bit pattern accumulator for the next state
    int steps;                     // steps taken
    int done;                      // 1: the time loop of this instance is over
    int status;                    // bit 0: non-finite wave speed detected
    int pad;
};

// Dynamic task tickets for the persistent launch (reset by the advance
// kernel after every step launch).
struct Sched {
    unsigned int ticket;
    unsigned int pad;
};

}  // namespace fvm2d
