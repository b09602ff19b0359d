// Host runtime behind the C ABI of include/fvm2d.h: device buffers, the
// device-resident time loop, launch geometry (with an autotuner), hipGraph
// capture of step chunks, and the plain / batched front ends.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <limits>
#include <string>
#include <vector>

#include "../../include/fvm2d.h"
#include "fvm2d_device.h"
#include "fvm2d_kernels.h"

namespace fvm2d {

static thread_local std::string g_last_error;

static int fail(int code, const std::string &msg) {
    g_last_error = msg;
    return code;
}

#define HIPCHK(expr)                                                                          \
    do {                                                                                      \
        hipError_t e_ = (expr);                                                               \
        if (e_ != hipSuccess)                                                                 \
            return fail(FVM2D_ERR_HIP, std::string(#expr) + ": " + hipGetErrorString(e_));    \
    } while (0)

static int check_params(int nx, int ny, const fvm2d_params *p) {
    if (!p) return fail(FVM2D_ERR_ARG, "params is null");
    if (nx < 6 || ny < 6)
        return fail(FVM2D_ERR_ARG, "mesh must be at least 6x6 cells (two ghost layers per side)");
    if ((long long)nx * ny > (1ll << 31) - 1)
        return fail(FVM2D_ERR_ARG, "mesh has more than 2^31-1 cells");
    if (!(p->width > 0.0) || !(p->height > 0.0)) return fail(FVM2D_ERR_ARG, "width/height must be > 0");
    if (!(p->cfl > 0.0)) return fail(FVM2D_ERR_ARG, "CFL must be > 0");
    if (!(p->gamma > 1.0)) return fail(FVM2D_ERR_ARG, "gamma must be > 1");
    if (std::isnan(p->t_end) || std::isnan(p->alpha)) return fail(FVM2D_ERR_ARG, "NaN parameter");
    if (p->strategy < 0 || p->strategy > 4) return fail(FVM2D_ERR_ARG, "unknown ghost-cell strategy");
    if (p->mode != FVM2D_MODE_FAST && p->mode != FVM2D_MODE_EXACT)
        return fail(FVM2D_ERR_ARG, "unknown arithmetic mode");
    return FVM2D_OK;
}

static int device_ok() {
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess || n < 1)
        return fail(FVM2D_ERR_NODEVICE, "no HIP device available (libfvm2d has no CPU fallback)");
    return FVM2D_OK;
}

// --------------------------------------------------------------------------
// Engine: n instances advanced by one step kernel per time step.
struct Engine {
    int n = 0;
    bool exact = false;
    std::vector<fvm2d_params> prm;
    std::vector<InstDesc> hd;        // host copies of the descriptors
    std::vector<long long> buf_off;  // offsets (doubles) of each instance's buffers in the pool
    double *pool = nullptr;
    long long pool_doubles = 0;
    InstDesc *d_desc = nullptr;
    Ctl *d_ctl = nullptr;
    Ctl *h_ctl = nullptr;            // pinned read-back
    Sched *d_sched = nullptr;
    int2 *d_tasks = nullptr;
    int ntasks = 0;
    int R = 16, wpb = 4, persistent = 0, grid = 0;
    int cu_count = 256;
    hipStream_t stream = nullptr;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    hipGraphExec_t gexec = nullptr;
    int gsteps = 0;                  // steps captured in gexec
    double *stage = nullptr;         // pinned staging for one instance
    long long stage_doubles = 0;
    std::vector<int> steps_total;
    bool begun = false;

    ~Engine() { release(); }

    void release() {
        if (gexec) hipGraphExecDestroy(gexec);
        gexec = nullptr;
        if (pool) hipFree(pool);
        if (d_desc) hipFree(d_desc);
        if (d_ctl) hipFree(d_ctl);
        if (d_sched) hipFree(d_sched);
        if (d_tasks) hipFree(d_tasks);
        if (h_ctl) hipHostFree(h_ctl);
        if (stage) hipHostFree(stage);
        if (e0) hipEventDestroy(e0);
        if (e1) hipEventDestroy(e1);
        if (stream) hipStreamDestroy(stream);
        pool = nullptr;
        d_desc = nullptr;
        d_ctl = nullptr;
        d_sched = nullptr;
        d_tasks = nullptr;
        h_ctl = nullptr;
        stage = nullptr;
        e0 = e1 = nullptr;
        stream = nullptr;
    }

    static void fill_desc(InstDesc &d, int nx, int ny, const fvm2d_params &p) {
        d.nx = nx;
        d.ny = ny;
        d.pitch = ((nx + 15) / 16) * 16;
        d.plane = (long long)d.pitch * ny;
        d.nw = (nx - 4 + kOwn - 1) / kOwn;
        d.strategy = p.strategy;
        d.max_steps = p.max_steps;
        d.task_base = 0;
        d.pad0 = 0;
        d.t_end = p.t_end;
        d.cfl = p.cfl;
        // h = std::min(width / Nx, height / Ny), src/wrappers.cuh:202
        d.hmin = std::min(p.width / double(nx), p.height / double(ny));
        // h = width / double(Nx) in MUSCL_rel_RP_GRP and forward_x_dir; the
        // Y-sweep runs them with (Ny, Nx, height), src/wrappers.cuh:219,223
        d.lx = make_line(p.width / double(nx), p.alpha);
        d.ly = make_line(p.height / double(ny), p.alpha);
        d.gas = make_gas(p.gamma);
    }

    int init(int count, const int *nx, const int *ny, const fvm2d_params *p) {
        int rc = device_ok();
        if (rc) return rc;
        if (count < 1) return fail(FVM2D_ERR_ARG, "instance count must be >= 1");
        n = count;
        prm.assign(p, p + count);
        exact = p[0].mode == FVM2D_MODE_EXACT;
        for (int i = 0; i < n; i++) {
            rc = check_params(nx[i], ny[i], &p[i]);
            if (rc) return rc;
            if ((p[i].mode == FVM2D_MODE_EXACT) != exact)
                return fail(FVM2D_ERR_ARG, "all instances of a batch must use the same mode");
        }
        hd.resize(n);
        buf_off.resize(n);
        steps_total.assign(n, 0);
        long long off = 0, maxplane = 0;
        for (int i = 0; i < n; i++) {
            fill_desc(hd[i], nx[i], ny[i], p[i]);
            buf_off[i] = off;
            off += 8 * hd[i].plane + 64;  // two buffers of four planes, 512-B separation
            off = (off + 63) / 64 * 64;
            maxplane = std::max(maxplane, hd[i].plane);
        }
        pool_doubles = off;
        HIPCHK(hipStreamCreateWithFlags(&stream, hipStreamNonBlocking));
        HIPCHK(hipEventCreate(&e0));
        HIPCHK(hipEventCreate(&e1));
        HIPCHK(hipMalloc(&pool, pool_doubles * sizeof(double)));
        HIPCHK(hipMemsetAsync(pool, 0, pool_doubles * sizeof(double), stream));
        for (int i = 0; i < n; i++) {
            hd[i].buf0 = pool + buf_off[i];
            hd[i].buf1 = hd[i].buf0 + 4 * hd[i].plane;
        }
        HIPCHK(hipMalloc(&d_desc, n * sizeof(InstDesc)));
        HIPCHK(hipMalloc(&d_ctl, n * sizeof(Ctl)));
        HIPCHK(hipMemsetAsync(d_ctl, 0, n * sizeof(Ctl), stream));
        HIPCHK(hipMalloc(&d_sched, sizeof(Sched)));
        HIPCHK(hipMemsetAsync(d_sched, 0, sizeof(Sched), stream));
        HIPCHK(hipHostMalloc(&h_ctl, n * sizeof(Ctl), hipHostMallocDefault));
        stage_doubles = 4 * maxplane;
        HIPCHK(hipHostMalloc(&stage, stage_doubles * sizeof(double), hipHostMallocDefault));
        int dev = 0;
        HIPCHK(hipGetDevice(&dev));
        hipDeviceProp_t prop;
        HIPCHK(hipGetDeviceProperties(&prop, dev));
        cu_count = prop.multiProcessorCount;
        R = default_strip_rows();
        return set_launch(R, 4, n > 1 ? 1 : 0);
    }

    int default_strip_rows() const {
        // aim at roughly two resident waves per SIMD over the whole batch
        long long cells = 0;
        for (const auto &d : hd) cells += (long long)(d.nx - 4) * (d.ny - 4);
        const double target_waves = 2.0 * 4 * cu_count;
        const double rows = cells / (double)kOwn / target_waves;
        int r = (int)std::lround(rows);
        return std::max(8, std::min(64, r));
    }

    int set_launch(int strip_rows, int waves_per_block, int persist) {
        if (strip_rows > 0) R = strip_rows;
        if (waves_per_block > 0) wpb = std::min(4, waves_per_block);
        if (persist >= 0) persistent = persist ? 1 : 0;
        if (R < 1) return fail(FVM2D_ERR_ARG, "strip rows must be >= 1");
        std::vector<int2> tasks;
        for (int i = 0; i < n; i++) {
            InstDesc &d = hd[i];
            d.R = R;
            d.ns = (d.ny - 4 + R - 1) / R;
            d.task_base = (int)tasks.size();
            const int cnt = d.nw * d.ns;
            for (int k = 0; k < cnt; k++) tasks.push_back(make_int2(i, k));
        }
        ntasks = (int)tasks.size();
        if (d_tasks) HIPCHK(hipFree(d_tasks));
        d_tasks = nullptr;
        HIPCHK(hipMalloc(&d_tasks, std::max(1, ntasks) * sizeof(int2)));
        HIPCHK(hipMemcpyAsync(d_tasks, tasks.data(), ntasks * sizeof(int2), hipMemcpyHostToDevice, stream));
        HIPCHK(hipMemcpyAsync(d_desc, hd.data(), n * sizeof(InstDesc), hipMemcpyHostToDevice, stream));
        const int blocks_needed = (ntasks + wpb - 1) / wpb;
        if (persistent) {
            int occ = step_kernel_occupancy(exact, wpb);
            if (occ < 1) occ = 1;
            grid = std::min(blocks_needed, occ * cu_count);
        } else {
            grid = blocks_needed;
        }
        if (grid < 1) grid = 1;
        drop_graph();
        HIPCHK(hipStreamSynchronize(stream));
        return FVM2D_OK;
    }

    void drop_graph() {
        if (gexec) hipGraphExecDestroy(gexec);
        gexec = nullptr;
        gsteps = 0;
    }

    int upload_desc() {
        HIPCHK(hipMemcpyAsync(d_desc, hd.data(), n * sizeof(InstDesc), hipMemcpyHostToDevice, stream));
        return FVM2D_OK;
    }

    // Host vec5d [ny][nx][5] -> current-state buffer (buf0) of instance i.
    int upload(int i, const double *U5) {
        if (i < 0 || i >= n || !U5) return fail(FVM2D_ERR_ARG, "bad instance or null state");
        const InstDesc &d = hd[i];
        HIPCHK(hipStreamSynchronize(stream));
        for (int k = 0; k < 4; k++)
            for (int j = 0; j < d.ny; j++) {
                double *dst = stage + k * d.plane + (long long)j * d.pitch;
                const double *s = U5 + ((long long)j * d.nx) * 5 + k;
                for (int c = 0; c < d.nx; c++) dst[c] = s[5 * c];
                for (int c = d.nx; c < d.pitch; c++) dst[c] = 0.0;
            }
        HIPCHK(hipMemcpyAsync(d.buf0, stage, 4 * d.plane * sizeof(double), hipMemcpyHostToDevice, stream));
        HIPCHK(hipStreamSynchronize(stream));
        begun = false;
        return FVM2D_OK;
    }

    int download(int i, double *U5) {
        if (i < 0 || i >= n || !U5) return fail(FVM2D_ERR_ARG, "bad instance or null state");
        const InstDesc &d = hd[i];
        HIPCHK(hipMemcpyAsync(stage, d.buf0, 4 * d.plane * sizeof(double), hipMemcpyDeviceToHost, stream));
        HIPCHK(hipStreamSynchronize(stream));
        for (int j = 0; j < d.ny; j++) {
            double *o = U5 + (long long)j * d.nx * 5;
            for (int c = 0; c < d.nx; c++) {
                for (int k = 0; k < 4; k++) o[5 * c + k] = stage[k * d.plane + (long long)j * d.pitch + c];
                o[5 * c + 4] = 0.0;
            }
        }
        return FVM2D_OK;
    }

    int planes_device(int i, double *d_planes, bool in) {
        if (i < 0 || i >= n || !d_planes) return fail(FVM2D_ERR_ARG, "bad instance or null pointer");
        const InstDesc &d = hd[i];
        for (int k = 0; k < 4; k++) {
            double *a = d.buf0 + k * d.plane;
            double *b = d_planes + (long long)k * d.nx * d.ny;
            if (in)
                HIPCHK(hipMemcpy2DAsync(a, d.pitch * sizeof(double), b, d.nx * sizeof(double),
                                        d.nx * sizeof(double), d.ny, hipMemcpyDeviceToDevice, stream));
            else
                HIPCHK(hipMemcpy2DAsync(b, d.nx * sizeof(double), a, d.pitch * sizeof(double),
                                        d.nx * sizeof(double), d.ny, hipMemcpyDeviceToDevice, stream));
        }
        HIPCHK(hipStreamSynchronize(stream));
        if (in) begun = false;
        return FVM2D_OK;
    }

    // The state of every instance is in buf0: reset its time loop to t = 0
    // and compute s_max of the initial state.
    int begin() {
        HIPCHK(hipMemsetAsync(d_ctl, 0, n * sizeof(Ctl), stream));
        for (int i = 0; i < n; i++)
            launch_speed(exact, d_desc, d_ctl, i, 0, (long long)hd[i].nx * hd[i].ny, stream);
        launch_ctl_start(d_desc, d_ctl, n, stream);
        HIPCHK(hipGetLastError());
        begun = true;
        return FVM2D_OK;
    }

    void launch_one() {
        launch_step(exact, d_desc, d_ctl, d_tasks, ntasks, d_sched, persistent, grid, wpb, stream);
    }

    int ensure_graph(int steps) {
        if (gexec && gsteps == steps) return FVM2D_OK;
        drop_graph();
        hipGraph_t g = nullptr;
        HIPCHK(hipStreamBeginCapture(stream, hipStreamCaptureModeThreadLocal));
        for (int k = 0; k < steps; k++) launch_one();
        HIPCHK(hipStreamEndCapture(stream, &g));
        hipError_t e = hipGraphInstantiate(&gexec, g, nullptr, nullptr, 0);
        hipGraphDestroy(g);
        if (e != hipSuccess) {
            gexec = nullptr;
            return fail(FVM2D_ERR_HIP, std::string("hipGraphInstantiate: ") + hipGetErrorString(e));
        }
        gsteps = steps;
        return FVM2D_OK;
    }

    static constexpr int kGraphSteps = 16;

    // Enqueue nsteps step launches (graph replays + a remainder of plain launches).
    int enqueue(int nsteps) {
        int rc = ensure_graph(kGraphSteps);
        if (rc) return rc;
        int k = 0;
        for (; k + kGraphSteps <= nsteps; k += kGraphSteps) HIPCHK(hipGraphLaunch(gexec, stream));
        for (; k < nsteps; k++) launch_one();
        HIPCHK(hipGetLastError());
        return FVM2D_OK;
    }

    int read_ctl() {
        HIPCHK(hipMemcpyAsync(h_ctl, d_ctl, n * sizeof(Ctl), hipMemcpyDeviceToHost, stream));
        HIPCHK(hipStreamSynchronize(stream));
        return FVM2D_OK;
    }

    bool all_done() const {
        for (int i = 0; i < n; i++)
            if (!h_ctl[i].done) return false;
        return true;
    }

    // Steps still needed by the slowest instance, estimated from its dt.
    int remaining_estimate() const {
        long long best = 0;
        for (int i = 0; i < n; i++) {
            const Ctl &c = h_ctl[i];
            if (c.done) continue;
            long long rem = 1ll << 30;
            if (prm[i].max_steps >= 0) rem = std::max(0, prm[i].max_steps - c.steps);
            if (std::isfinite(prm[i].t_end) && c.dt > 0.0) {
                const double r = (prm[i].t_end - c.t) / c.dt;
                rem = std::min(rem, (long long)std::ceil(r * 1.02) + 2);
            }
            best = std::max(best, rem);
        }
        return (int)std::min<long long>(best, 1 << 20);
    }

    // Device-resident time loop until every instance is done.
    int run_to_completion(double *ms_out) {
        if (!begun) {
            int rc = begin();
            if (rc) return rc;
        }
        HIPCHK(hipEventRecord(e0, stream));
        int rc = read_ctl();
        if (rc) return rc;
        int chunk = 8;
        while (!all_done()) {
            rc = enqueue(chunk);
            if (rc) return rc;
            rc = read_ctl();
            if (rc) return rc;
            chunk = std::max(1, remaining_estimate());
        }
        HIPCHK(hipEventRecord(e1, stream));
        HIPCHK(hipEventSynchronize(e1));
        float ms = 0.f;
        HIPCHK(hipEventElapsedTime(&ms, e0, e1));
        if (ms_out) *ms_out = ms;
        return finish_run();
    }

    // Time every candidate launch geometry on a scratch copy of the states
    // (t_end lifted so no instance stops) and keep the fastest.
    int autotune(int steps_per_trial, double *best_ms_out) {
        if (steps_per_trial < 1) steps_per_trial = 4;
        long long save_doubles = 0;
        for (int i = 0; i < n; i++) save_doubles += 4 * hd[i].plane;
        double *save = nullptr;
        HIPCHK(hipMalloc(&save, save_doubles * sizeof(double)));
        auto copy_states = [&](bool to_save) -> int {
            long long o = 0;
            for (int i = 0; i < n; i++) {
                const long long bytes = 4 * hd[i].plane * sizeof(double);
                if (to_save) HIPCHK(hipMemcpyAsync(save + o, hd[i].buf0, bytes, hipMemcpyDeviceToDevice, stream));
                else HIPCHK(hipMemcpyAsync(hd[i].buf0, save + o, bytes, hipMemcpyDeviceToDevice, stream));
                o += 4 * hd[i].plane;
            }
            return FVM2D_OK;
        };
        int rc = copy_states(true);
        std::vector<double> tend(n);
        std::vector<int> msteps(n);
        for (int i = 0; i < n; i++) {
            tend[i] = hd[i].t_end;
            msteps[i] = hd[i].max_steps;
            hd[i].t_end = std::numeric_limits<double>::infinity();
            hd[i].max_steps = -1;
        }
        int maxny = 0;
        for (int i = 0; i < n; i++) maxny = std::max(maxny, hd[i].ny - 4);
        const int Rs[] = {4, 8, 12, 16, 24, 32, 48, 64};
        double best = std::numeric_limits<double>::infinity();
        int bR = R, bW = wpb, bP = persistent;
        for (int r : Rs) {
            if (rc) break;
            if (r > std::max(4, maxny)) continue;
            for (int w : {1, 2, 4}) {
                for (int pers : {0, 1}) {
                    rc = set_launch(r, w, pers);
                    if (!rc) rc = copy_states(false);
                    if (!rc) rc = begin();
                    if (!rc) rc = enqueue(2);
                    if (rc) break;
                    hipEventRecord(e0, stream);
                    rc = enqueue(steps_per_trial);
                    if (rc) break;
                    hipEventRecord(e1, stream);
                    if (hipEventSynchronize(e1) != hipSuccess) {
                        rc = fail(FVM2D_ERR_HIP, "autotune trial failed");
                        break;
                    }
                    float ms = 0.f;
                    hipEventElapsedTime(&ms, e0, e1);
                    const double per = ms / steps_per_trial;
                    if (per < best) {
                        best = per;
                        bR = r;
                        bW = w;
                        bP = pers;
                    }
                }
                if (rc) break;
            }
        }
        for (int i = 0; i < n; i++) {
            hd[i].t_end = tend[i];
            hd[i].max_steps = msteps[i];
        }
        int rc2 = set_launch(bR, bW, bP);
        if (!rc2) rc2 = copy_states(false);
        hipStreamSynchronize(stream);
        hipFree(save);
        begun = false;
        if (best_ms_out) *best_ms_out = best;
        return rc ? rc : rc2;
    }

    // After a run: move each final state to buf0 (swap buffers when the
    // step count is odd) so that the next run continues from it.
    int finish_run() {
        int rc = read_ctl();
        if (rc) return rc;
        int status = 0;
        for (int i = 0; i < n; i++) {
            if (h_ctl[i].steps & 1) std::swap(hd[i].buf0, hd[i].buf1);
            steps_total[i] += h_ctl[i].steps;
            status |= h_ctl[i].status;
        }
        rc = upload_desc();
        if (rc) return rc;
        HIPCHK(hipStreamSynchronize(stream));
        begun = false;
        if (status & 1) return fail(FVM2D_ERR_NONFINITE, "non-finite wave speed (NaN state) reached");
        return FVM2D_OK;
    }
};

}  // namespace fvm2d

using namespace fvm2d;

struct fvm2d_solver {
    Engine eng;
    int nx = 0, ny = 0;
};

struct fvm2d_batch {
    Engine eng;
};

extern "C" {

void fvm2d_params_default(fvm2d_params *p) {
    if (!p) return;
    p->width = 1.0;
    p->height = 1.0;
    p->t_end = 0.0;
    p->cfl = 0.5;
    p->alpha = 1.9;
    p->gamma = 1.4;
    p->strategy = FVM2D_OUTFLOW;
    p->max_steps = -1;
    p->mode = FVM2D_MODE_FAST;
    p->reserved = 0;
}

const char *fvm2d_last_error(void) { return g_last_error.c_str(); }

int fvm2d_device_available(void) {
    int n = 0;
    return (hipGetDeviceCount(&n) == hipSuccess && n > 0) ? 1 : 0;
}

int fvm2d_create(int nx, int ny, const fvm2d_params *p, fvm2d_solver **out) {
    if (!out) return fail(FVM2D_ERR_ARG, "out is null");
    *out = nullptr;
    int rc = check_params(nx, ny, p);
    if (rc) return rc;
    auto *s = new fvm2d_solver();
    s->nx = nx;
    s->ny = ny;
    rc = s->eng.init(1, &nx, &ny, p);
    if (rc) {
        delete s;
        return rc;
    }
    *out = s;
    return FVM2D_OK;
}

int fvm2d_set_params(fvm2d_solver *s, const fvm2d_params *p) {
    if (!s) return fail(FVM2D_ERR_ARG, "solver is null");
    int rc = check_params(s->nx, s->ny, p);
    if (rc) return rc;
    if ((p->mode == FVM2D_MODE_EXACT) != s->eng.exact) {
        s->eng.exact = p->mode == FVM2D_MODE_EXACT;
        s->eng.drop_graph();
    }
    s->eng.prm[0] = *p;
    double *b0 = s->eng.hd[0].buf0, *b1 = s->eng.hd[0].buf1;
    Engine::fill_desc(s->eng.hd[0], s->nx, s->ny, *p);
    s->eng.hd[0].buf0 = b0;
    s->eng.hd[0].buf1 = b1;
    return s->eng.set_launch(s->eng.R, s->eng.wpb, s->eng.persistent);
}

int fvm2d_upload(fvm2d_solver *s, const double *U5) {
    if (!s) return fail(FVM2D_ERR_ARG, "solver is null");
    return s->eng.upload(0, U5);
}

int fvm2d_upload_planes_device(fvm2d_solver *s, const double *d_planes) {
    if (!s) return fail(FVM2D_ERR_ARG, "solver is null");
    return s->eng.planes_device(0, const_cast<double *>(d_planes), true);
}

int fvm2d_download_planes_device(fvm2d_solver *s, double *d_planes) {
    if (!s) return fail(FVM2D_ERR_ARG, "solver is null");
    return s->eng.planes_device(0, d_planes, false);
}

int fvm2d_run(fvm2d_solver *s, int *steps_out, double *t_out, double *ms_out) {
    if (!s) return fail(FVM2D_ERR_ARG, "solver is null");
    s->eng.begun = false;
    double ms = 0.0;
    int rc = s->eng.run_to_completion(&ms);
    if (steps_out) *steps_out = s->eng.h_ctl[0].steps;
    if (t_out) *t_out = s->eng.h_ctl[0].t;
    if (ms_out) *ms_out = ms;
    return rc;
}

int fvm2d_download(fvm2d_solver *s, double *U5) {
    if (!s) return fail(FVM2D_ERR_ARG, "solver is null");
    return s->eng.download(0, U5);
}

int fvm2d_steps_total(const fvm2d_solver *s) { return s ? s->eng.steps_total[0] : 0; }

int fvm2d_destroy(fvm2d_solver *s) {
    delete s;
    return FVM2D_OK;
}

int fvm2d_solve(double *U5, int nx, int ny, const fvm2d_params *p, int *steps_out, double *ms_out) {
    if (!U5) return fail(FVM2D_ERR_ARG, "state is null");
    fvm2d_solver *s = nullptr;
    int rc = fvm2d_create(nx, ny, p, &s);
    if (rc) return rc;
    rc = fvm2d_upload(s, U5);
    if (!rc) rc = fvm2d_run(s, steps_out, nullptr, ms_out);
    int rc2 = fvm2d_download(s, U5);
    fvm2d_destroy(s);
    return rc ? rc : rc2;
}

int fvm2d_set_launch(fvm2d_solver *s, int strip_rows, int waves_per_block, int persistent) {
    if (!s) return fail(FVM2D_ERR_ARG, "solver is null");
    return s->eng.set_launch(strip_rows, waves_per_block, persistent);
}

int fvm2d_get_launch(const fvm2d_solver *s, int *strip_rows, int *waves_per_block, int *persistent) {
    if (!s) return fail(FVM2D_ERR_ARG, "solver is null");
    if (strip_rows) *strip_rows = s->eng.R;
    if (waves_per_block) *waves_per_block = s->eng.wpb;
    if (persistent) *persistent = s->eng.persistent;
    return FVM2D_OK;
}

int fvm2d_begin(fvm2d_solver *s) {
    if (!s) return fail(FVM2D_ERR_ARG, "solver is null");
    return s->eng.begin();
}

int fvm2d_advance(fvm2d_solver *s, int nsteps) {
    if (!s) return fail(FVM2D_ERR_ARG, "solver is null");
    if (!s->eng.begun) {
        int rc = s->eng.begin();
        if (rc) return rc;
    }
    return s->eng.enqueue(nsteps);
}

int fvm2d_sync(fvm2d_solver *s, int *steps_out, double *t_out, double *dt_out) {
    if (!s) return fail(FVM2D_ERR_ARG, "solver is null");
    int rc = s->eng.read_ctl();
    if (rc) return rc;
    if (steps_out) *steps_out = s->eng.h_ctl[0].steps;
    if (t_out) *t_out = s->eng.h_ctl[0].t;
    if (dt_out) *dt_out = s->eng.h_ctl[0].dt;
    if (s->eng.h_ctl[0].status & 1) return fail(FVM2D_ERR_NONFINITE, "non-finite wave speed (NaN state) reached");
    return FVM2D_OK;
}

int fvm2d_time_steps(fvm2d_solver *s, int nsteps, double *ms_out) {
    if (!s) return fail(FVM2D_ERR_ARG, "solver is null");
    Engine &E = s->eng;
    if (!E.begun) {
        int rc = E.begin();
        if (rc) return rc;
    }
    int rc = E.ensure_graph(Engine::kGraphSteps);
    if (rc) return rc;
    HIPCHK(hipEventRecord(E.e0, E.stream));
    rc = E.enqueue(nsteps);
    if (rc) return rc;
    HIPCHK(hipEventRecord(E.e1, E.stream));
    HIPCHK(hipEventSynchronize(E.e1));
    float ms = 0.f;
    HIPCHK(hipEventElapsedTime(&ms, E.e0, E.e1));
    if (ms_out) *ms_out = ms;
    return FVM2D_OK;
}

int fvm2d_autotune(fvm2d_solver *s, int steps_per_trial, double *best_ms_out) {
    if (!s) return fail(FVM2D_ERR_ARG, "solver is null");
    return s->eng.autotune(steps_per_trial, best_ms_out);
}

int fvm2d_batch_create(int n, const int *nx, const int *ny, const fvm2d_params *p,
                       fvm2d_batch **out) {
    if (!out || !nx || !ny || !p) return fail(FVM2D_ERR_ARG, "null argument");
    *out = nullptr;
    auto *b = new fvm2d_batch();
    int rc = b->eng.init(n, nx, ny, p);
    if (rc) {
        delete b;
        return rc;
    }
    *out = b;
    return FVM2D_OK;
}

int fvm2d_batch_upload(fvm2d_batch *b, int i, const double *U5) {
    if (!b) return fail(FVM2D_ERR_ARG, "batch is null");
    return b->eng.upload(i, U5);
}

int fvm2d_batch_begin(fvm2d_batch *b) {
    if (!b) return fail(FVM2D_ERR_ARG, "batch is null");
    return b->eng.begin();
}

int fvm2d_batch_run(fvm2d_batch *b, int nsteps, double *ms_out) {
    if (!b) return fail(FVM2D_ERR_ARG, "batch is null");
    Engine &E = b->eng;
    if (!E.begun) {
        int rc = E.begin();
        if (rc) return rc;
    }
    int rc = E.ensure_graph(Engine::kGraphSteps);
    if (rc) return rc;
    HIPCHK(hipEventRecord(E.e0, E.stream));
    rc = E.enqueue(nsteps);
    if (rc) return rc;
    HIPCHK(hipEventRecord(E.e1, E.stream));
    HIPCHK(hipEventSynchronize(E.e1));
    float ms = 0.f;
    HIPCHK(hipEventElapsedTime(&ms, E.e0, E.e1));
    if (ms_out) *ms_out = ms;
    rc = E.read_ctl();
    if (rc) return rc;
    for (int i = 0; i < E.n; i++)
        if (E.h_ctl[i].status & 1)
            return fail(FVM2D_ERR_NONFINITE, "instance " + std::to_string(i) + ": non-finite wave speed");
    return FVM2D_OK;
}

int fvm2d_batch_status(fvm2d_batch *b, int i, int *steps_out, double *t_out) {
    if (!b) return fail(FVM2D_ERR_ARG, "batch is null");
    if (i < 0 || i >= b->eng.n) return fail(FVM2D_ERR_ARG, "bad instance");
    int rc = b->eng.read_ctl();
    if (rc) return rc;
    if (steps_out) *steps_out = b->eng.h_ctl[i].steps;
    if (t_out) *t_out = b->eng.h_ctl[i].t;
    return FVM2D_OK;
}

int fvm2d_batch_download(fvm2d_batch *b, int i, double *U5) {
    if (!b) return fail(FVM2D_ERR_ARG, "batch is null");
    Engine &E = b->eng;
    if (i < 0 || i >= E.n) return fail(FVM2D_ERR_ARG, "bad instance");
    int rc = E.read_ctl();
    if (rc) return rc;
    // the current state of instance i is buf[steps & 1]
    const InstDesc &d = E.hd[i];
    double *cur = (E.h_ctl[i].steps & 1) ? d.buf1 : d.buf0;
    HIPCHK(hipMemcpyAsync(E.stage, cur, 4 * d.plane * sizeof(double), hipMemcpyDeviceToHost, E.stream));
    HIPCHK(hipStreamSynchronize(E.stream));
    for (int j = 0; j < d.ny; j++) {
        double *o = U5 + (long long)j * d.nx * 5;
        for (int c = 0; c < d.nx; c++) {
            for (int k = 0; k < 4; k++) o[5 * c + k] = E.stage[k * d.plane + (long long)j * d.pitch + c];
            o[5 * c + 4] = 0.0;
        }
    }
    return FVM2D_OK;
}

int fvm2d_batch_set_launch(fvm2d_batch *b, int strip_rows, int waves_per_block, int persistent) {
    if (!b) return fail(FVM2D_ERR_ARG, "batch is null");
    return b->eng.set_launch(strip_rows, waves_per_block, persistent);
}

int fvm2d_batch_autotune(fvm2d_batch *b, int steps_per_trial, double *best_ms_out) {
    if (!b) return fail(FVM2D_ERR_ARG, "batch is null");
    return b->eng.autotune(steps_per_trial, best_ms_out);
}

int fvm2d_batch_destroy(fvm2d_batch *b) {
    delete b;
    return FVM2D_OK;
}

}  // extern "C"
