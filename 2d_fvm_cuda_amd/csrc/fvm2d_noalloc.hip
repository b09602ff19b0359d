// fvm2d_solve_device_noalloc: the device-planes solve without any device
// memory allocation, for callers that cannot allocate (code running under
// stream capture setup, tight serving loops, harnesses that forbid
// hipMalloc).  Everything the time loop needs beyond the caller's state --
// the second time-loop buffer, the mesh descriptor, its control block, the
// task ticket and the task list -- lives in one
// file-scope device array.  Each GPU holds its own copy of that array for
// the life of the process; host code reaches it only through the address
// hipGetSymbolAddress returns for the current device.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <mutex>

#include "fvm2d_engine.h"

namespace fvm2d {

constexpr long long kNoAllocMaxCells = 1ll << 20;  // e.g. 1024 x 1024 with ghosts
constexpr int kNoAllocMaxTasks = 16384;
constexpr size_t kAlign = 256;

constexpr size_t align_up(size_t x) { return (x + kAlign - 1) / kAlign * kAlign; }

// Layout of the array: [state buffer][InstDesc][Ctl][Sched][tasks]
constexpr size_t kOffState = 0;
constexpr size_t kOffDesc = align_up(kOffState + 4 * (size_t)kNoAllocMaxCells * sizeof(double));
constexpr size_t kOffCtl = align_up(kOffDesc + sizeof(InstDesc));
constexpr size_t kOffSched = align_up(kOffCtl + sizeof(Ctl));
constexpr size_t kOffTasks = align_up(kOffSched + sizeof(Sched));
constexpr size_t kScratchBytes = align_up(kOffTasks + kNoAllocMaxTasks * sizeof(int4));

static __device__ __align__(256) unsigned char scratch[kScratchBytes];

// Device address of scratch per device (0 until looked up).
constexpr int kMaxDevices = 64;
static unsigned char *g_scratch[kMaxDevices] = {};
static std::mutex g_scratch_mu;

static int scratch_ptr(unsigned char **out) {
    int dev = 0;
    HIPCHK(hipGetDevice(&dev));
    if (dev < 0 || dev >= kMaxDevices) return fail(FVM2D_ERR_ARG, "device index out of range");
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (!g_scratch[dev]) {
        void *p = nullptr;
        HIPCHK(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
        g_scratch[dev] = static_cast<unsigned char *>(p);
    }
    *out = g_scratch[dev];
    return FVM2D_OK;
}

// Writes the descriptor, clears the control block and the task ticket (a
// kernel, not a memset), and lays out the strips: task k is
// window k mod nw, rows [2 + s R, min(ny - 2, 2 + (s + 1) R)) with s = k / nw.
__global__ void __launch_bounds__(256) noalloc_init_kernel(InstDesc d, InstDesc *desc, Ctl *ctl,
                                                           Sched *sched, int4 *tasks, int ntasks) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k == 0) {
        *desc = d;
        *ctl = Ctl{};
        *sched = Sched{0u, 0u};
    }
    if (k < ntasks) {
        const int w = k % d.nw, s = k / d.nw;
        const int r0 = 2 + s * d.R;
        tasks[k] = make_int4(0, w, r0, min(d.ny - 2, r0 + d.R));
    }
}

}  // namespace fvm2d

using namespace fvm2d;

extern "C" int fvm2d_solve_device_noalloc(double *d_planes, int nx, int ny, const fvm2d_params *p,
                                          int *steps_out, double *t_out) {
    if (!d_planes) return fail(FVM2D_ERR_ARG, "state is null");
    int rc = check_params(nx, ny, p);
    if (rc) return rc;
    if ((long long)nx * ny > kNoAllocMaxCells)
        return fail(FVM2D_ERR_ARG, "mesh larger than the no-allocation limit of 1048576 cells");
    if ((rc = device_ok())) return rc;
    unsigned char *base = nullptr;
    if ((rc = scratch_ptr(&base))) return rc;

    InstDesc d;
    Engine::fill_desc(d, nx, ny, *p);
    d.pitch = nx;  // the caller's planes are contiguous [4][ny][nx]
    d.plane = (long long)nx * ny;
    d.buf0 = d_planes;
    d.buf1 = reinterpret_cast<double *>(base + kOffState);
    // strips of R rows over the nw windows, R the lowest height (>= 8)
    // whose task count fits the task array
    const int H = ny - 4;
    int R = 8;
    while ((long long)d.nw * ((H + R - 1) / R) > kNoAllocMaxTasks) R *= 2;
    R = std::min(R, std::max(H, 1));
    const int ntasks = d.nw * ((H + R - 1) / R);
    d.R = R;
    d.task_base = 0;
    d.ntask = ntasks;

    InstDesc *d_desc = reinterpret_cast<InstDesc *>(base + kOffDesc);
    Ctl *d_ctl = reinterpret_cast<Ctl *>(base + kOffCtl);
    Sched *d_sched = reinterpret_cast<Sched *>(base + kOffSched);
    int4 *d_tasks = reinterpret_cast<int4 *>(base + kOffTasks);
    const bool exact = p->mode == FVM2D_MODE_EXACT;
    hipStream_t st = nullptr;
    hipLaunchKernelGGL(noalloc_init_kernel, dim3((ntasks + 255) / 256), dim3(256), 0, st, d, d_desc, d_ctl,
                       d_sched, d_tasks, ntasks);
    launch_speed(exact, d_desc, d_ctl, 0, 0, ny, st);
    launch_ctl_start(d_desc, d_ctl, 1, nullptr, st);
    HIPCHK(hipGetLastError());
    const int grid = (ntasks + 3) / 4;
    Ctl c;
    for (;;) {
        HIPCHK(hipMemcpy(&c, d_ctl, sizeof(Ctl), hipMemcpyDeviceToHost));
        if (c.done) break;
        // launches up to the estimated end; a launch after it exits at once
        long long rem = 1ll << 20;
        if (p->max_steps >= 0) rem = std::max(0, p->max_steps - c.steps);
        if (c.dt > 0.0 && p->t_end < 1e300) {
            const double r = (p->t_end - c.t) / c.dt;
            if (r < 1e9) rem = std::min(rem, (long long)(r * 1.02) + 2);
        }
        const int chunk = (int)std::max(1ll, std::min(rem, 4096ll));
        for (int k = 0; k < chunk; k++) launch_step(exact, d_desc, d_ctl, 1, d_tasks, ntasks, d_sched, 0, grid, 4, st);
        HIPCHK(hipGetLastError());
    }
    // the final state is buf[steps & 1]
    if (c.steps & 1)
        HIPCHK(hipMemcpy(d_planes, d.buf1, 4 * (size_t)d.plane * sizeof(double), hipMemcpyDeviceToDevice));
    if (steps_out) *steps_out = c.steps;
    if (t_out) *t_out = c.t;
    if (c.status & 1) return fail(FVM2D_ERR_NONFINITE, "non-finite wave speed (NaN state) reached");
    return FVM2D_OK;
}
