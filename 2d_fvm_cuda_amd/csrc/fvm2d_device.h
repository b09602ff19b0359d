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

// Arrival slots of the step kernel's end-of-task reduction.  The tasks of
// an instance are spread over kCtlSlots slots (task index mod kCtlSlots),
// each on its own 128-byte line; the last task of a slot forwards the slot's
// maximum to the instance's Ctl.  Without the slots every wave of a launch
// (about 2000 at 2000 x 2000) would meet on one address at the end.
constexpr int kCtlSlots = 32;
struct alignas(128) CtlSlot {
    unsigned long long smax;  // bit-pattern maximum of the slot's tasks
    unsigned int count;       // tasks of the slot finished in this launch
    unsigned int pad[29];
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
    CtlSlot *slots;    // kCtlSlots arrival slots, zero between launches
};

// Per-instance time-loop control, resident on the device.
struct Ctl {
    double t;                      // current time
    double dt;                     // step size for the next step
    double dt_used;                // step size of the last step taken
    double smax;                   // max |(u,v)| + c over the current state
    unsigned long long smax_next;  // bit pattern accumulator for the next state
    unsigned int waves_done;       // tasks of this instance finished in this launch
    int steps;                     // steps taken
    int done;                      // 1: the time loop of this instance is over
    int status;                    // bit 0: non-finite wave speed detected
};

// Dynamic task tickets for the persistent launch.
struct Sched {
    unsigned int ticket;
    unsigned int exited;
};

}  // namespace fvm2d
