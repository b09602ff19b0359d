// Host-side engine behind the C ABI: device buffers of n instances, the
// device-resident time loop, launch geometry (with an autotuner) and hipGraph
// capture of step chunks.  Shared by the forward and adjoint front ends.
#pragma once

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <limits>
#include <string>
#include <thread>
#include <vector>

#include "../../include/fvm2d.h"
#include "fvm2d_device.h"
#include "fvm2d_kernels.h"

namespace fvm2d {

inline thread_local std::string g_last_error;

inline int fail(int code, const std::string &msg) {
    g_last_error = msg;
    return code;
}

#define HIPCHK(expr)                                                                          \
    do {                                                                                      \
        hipError_t e_ = (expr);                                                               \
        if (e_ != hipSuccess)                                                                 \
            return fail(FVM2D_ERR_HIP, std::string(#expr) + ": " + hipGetErrorString(e_));    \
    } while (0)

inline int check_params(int nx, int ny, const fvm2d_params *p) {
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

inline int device_ok() {
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess || n < 1)
        return fail(FVM2D_ERR_NODEVICE, "no HIP device available (libfvm2d has no CPU fallback)");
    return FVM2D_OK;
}

// --------------------------------------------------------------------------
// Row weight of the older waves' strips in a one-round launch of one mesh
// (see set_launch).
constexpr double kOlderWeight = 2.0;

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
    double *d_t0 = nullptr, *h_t0 = nullptr;  // per-instance start times
    int4 *d_tasks = nullptr;
    std::vector<int4> h_tasks;       // host copy of d_tasks (source of its async upload)
    int ntasks = 0;
    int R = 16, wpb = 4, persistent = 0, grid = 0;
    int cu_count = 256;
    hipStream_t stream = nullptr;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    hipGraphExec_t gexec = nullptr;
    int gsteps = 0;                  // steps captured in gexec
    double *stage = nullptr;         // pinned staging for one instance
    double *d_frame = nullptr;       // record-mode frame (4 planes of the largest mesh)
    unsigned char *d_gray = nullptr; // its 8-bit density image
    unsigned long long *d_range = nullptr;  // the image's min/max keys
    long long stage_doubles = 0;
    std::vector<int> steps_total;
    bool begun = false;

    ~Engine() { release(); }

    void release() {
        // work still queued on the stream (launches enqueued without a sync)
        // may use every buffer and the graph below
        if (stream) hipStreamSynchronize(stream);
        if (gexec) hipGraphExecDestroy(gexec);
        gexec = nullptr;
        if (pool) hipFree(pool);
        if (d_desc) hipFree(d_desc);
        if (d_ctl) hipFree(d_ctl);
        if (d_sched) hipFree(d_sched);
        if (d_tasks) hipFree(d_tasks);
        if (h_ctl) hipHostFree(h_ctl);
        if (d_t0) hipFree(d_t0);
        if (h_t0) hipHostFree(h_t0);
        d_t0 = nullptr;
        h_t0 = nullptr;
        if (stage) hipHostFree(stage);
        if (d_frame) hipFree(d_frame);
        d_frame = nullptr;
        if (d_gray) hipFree(d_gray);
        d_gray = nullptr;
        if (d_range) hipFree(d_range);
        d_range = nullptr;
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
        HIPCHK(hipMalloc(&d_t0, n * sizeof(double)));
        HIPCHK(hipHostMalloc(&h_t0, n * sizeof(double), hipHostMallocDefault));
        stage_doubles = 4 * maxplane;
        HIPCHK(hipHostMalloc(&stage, stage_doubles * sizeof(double), hipHostMallocDefault));
        int dev = 0;
        HIPCHK(hipGetDevice(&dev));
        hipDeviceProp_t prop;
        HIPCHK(hipGetDeviceProperties(&prop, dev));
        cu_count = prop.multiProcessorCount;
        if (n == 1) return set_launch(one_mesh_rows(), 4, 0);
        R = default_strip_rows();
        return set_launch(R, 4, 1);
    }

    // Strip height of one mesh launched without tuning (four waves per
    // block, one wave per task): the lowest strips whose count over all
    // column windows still fits the fewest rounds of resident waves that
    // keep strips at most 96 rows high.  mesh_strips then cuts them so that
    // one or two rounds are filled exactly.  At 2000 x 2000 on 256 CUs this
    // is one round at R = 34, the geometry the autotuner picks.  A strip
    // also pays a fixed cost (its halo rows and the march's start), so when
    // one round would give strips below kMinOneMeshRows rows, half a round
    // (one wave per SIMD) of twice as high strips is used, at least 4 rows
    // (measured per step, against one round and against the sixteenth of a
    // round of 11-row strips an earlier rule went down to: 256^2 21.5 us at
    // R = 4 against 39.9 us at R = 11, 512^2 30.6 us at R = 5 against 38.1
    // us, 640^2 35.3 us at R = 7 against 41.7 us for one round, 768^2 42.3
    // us at R = 10 against 48.8 us, 1000^2 59.1 us at R = 17 against 59.3
    // us; from 1100^2 on one round is faster, 1200^2 66.8 us at R = 12
    // against 78.1 us at R = 24).
    static constexpr long long kMinOneMeshRows = 10;
    int one_mesh_rows() const {
        const InstDesc &d = hd[0];
        const long long H = d.ny - 4;
        const long long slots = (long long)std::max(1, step_kernel_occupancy(exact, 4)) * 4 * cu_count;
        auto rows_for = [&](long long per_window) { return (H + per_window - 1) / per_window; };
        for (long long k = 1; k <= 64; k++) {
            const long long per_window = k * slots / d.nw;
            if (per_window < 1) continue;
            const long long r = rows_for(per_window);
            if (r > 96) continue;
            if (k == 1 && r < kMinOneMeshRows) {
                const long long pw = std::max(1ll, slots / 2 / d.nw);
                return (int)std::min(96ll, std::max(4ll, rows_for(pw)));
            }
            return (int)std::max<long long>(4, r);
        }
        return 64;
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

    // Strips of one mesh (instance inst: nw column windows, ny rows with
    // ghosts) for strip height R, appended to tasks as {inst, window, first
    // row, end row}.  single_fixed: the launch holds this mesh alone with
    // one wave per task; slots: resident waves of the launch; S: SIMDs
    // (4 * CUs).  No device state is used.
    static void mesh_strips(int inst, int nw, int ny, int R, bool single_fixed, long long slots,
                            long long S, int wpb, std::vector<int4> &tasks) {
        const int i = inst;
        const int H = ny - 4;
        const int ns = (H + R - 1) / R;
        // One mesh whose strips of R rows leave the last round of
        // resident waves partly empty (at 2000 x 2000 and R = 34: 2006
        // tasks for 2048 wave slots, so ten CUs run half their waves):
        // cut the windows into more, slightly lower strips so that the
        // task count fills the round exactly.  Windows 0, nw-1, 1,
        // nw-2, ... take the extra strips (edge windows also write the
        // ghost cells).  The cut decides only which wave computes
        // which rows.
        const long long T0 = (long long)nw * ns;
        const long long k = (T0 + slots - 1) / slots;
        const long long T = k * slots;
        if (single_fixed && k <= 2 && T0 < T && T0 * 10 >= T * 8 &&
            T <= (long long)nw * (H / 4)) {
            const int base = (int)(T / nw), extra = (int)(T % nw);
            std::vector<int> nsw(nw, base);
            for (int e = 0; e < extra; e++) nsw[(e & 1) ? nw - 1 - e / 2 : e / 2]++;
            // Strip heights.  In one round of two waves per SIMD
            // launched four waves per block (the dispatch measured; two
            // per block place the second round otherwise), the
            // first S tasks are the older waves, which take
            // most issue slots while both run; the SIMD then runs its
            // younger wave alone at a lower issue rate.  Giving the
            // older waves' strips kOlderWeight times the rows of the
            // younger waves' strips keeps both busy for longer
            // (2000 x 2000 over 1500 steps: weight 1.6-2.4 all 2.2% less
            // time per step than equal strips, 3.0 worse).  older[w]:
            // strips of window w among the first S tasks.
            std::vector<int> older(nw, 0);
            if (T == 2 * S && wpb == 4) {
                long long idx = 0;
                for (int s = 0; s <= base && idx < S; s++)
                    for (int w = 0; w < nw && idx < S; w++)
                        if (s < nsw[w]) older[w]++, idx++;
            }
            auto bound = [&](int w, int s) {
                const int m = nsw[w], a = older[w];
                const double tot = a * kOlderWeight + (m - a);
                const double cum = s <= a ? s * kOlderWeight : a * kOlderWeight + (s - a);
                return 2 + (int)std::llround(H * cum / tot);
            };
            for (int s = 0; s <= base; s++)
                for (int w = 0; w < nw; w++) {
                    if (s >= nsw[w]) continue;
                    tasks.push_back(make_int4(i, w, bound(w, s), bound(w, s + 1)));
                }
        } else {
            // ns strips of at most R rows, heights differing by at most one
            // (not ns - 1 strips of R rows and a short last one)
            for (int s = 0; s < ns; s++)
                for (int w = 0; w < nw; w++) {
                    const int r0 = 2 + (int)((long long)s * H / ns);
                    const int r1 = 2 + (int)((long long)(s + 1) * H / ns);
                    tasks.push_back(make_int4(i, w, r0, r1));
                }
        }
    }

    // The second round of one mesh's tasks in reverse order (see set_launch).
    static void mirror_second_round(std::vector<int4> &tasks, long long S) {
        const long long nt = (long long)tasks.size();
        if (nt > S && nt <= 2 * S) std::reverse(tasks.begin() + S, tasks.end());
    }

    // Change the launch geometry.  On any failure the current geometry (task
    // list, descriptors, grid and graph) is left as it was.
    int set_launch(int strip_rows, int waves_per_block, int persist) {
        const int newR = strip_rows > 0 ? strip_rows : R;
        const int newW = waves_per_block > 0 ? std::min(4, waves_per_block) : wpb;
        const int newP = persist >= 0 ? (persist ? 1 : 0) : persistent;
        {
            // task positions are ints: refuse geometries with more tasks
            // (the round-filling cut at most doubles a mesh's count)
            long long total = 0;
            for (int i = 0; i < n; i++)
                total += 2LL * hd[i].nw * ((hd[i].ny - 4 + newR - 1) / newR) + 1;
            if (total > (1LL << 30)) return fail(FVM2D_ERR_ARG, "strip rows too low for this mesh");
        }
        std::vector<int4> tasks;
        std::vector<InstDesc> nd = hd;
        // resident waves of the launch: blocks per CU x waves per block x CUs
        const int occ = std::max(1, step_kernel_occupancy(exact, newW));
        const long long slots = (long long)occ * newW * cu_count;
        for (int i = 0; i < n; i++) {
            InstDesc &d = nd[i];
            d.R = newR;
            d.task_base = (int)tasks.size();
            mesh_strips(i, d.nw, d.ny, newR, n == 1 && !newP, slots, 4LL * cu_count, newW, tasks);
            d.ntask = (int)tasks.size() - d.task_base;
        }
        const int nt = (int)tasks.size();
        // Task order of one mesh whose tasks fill one round of two waves per
        // SIMD: the dispatcher gives the first 4 * cu_count positions one
        // SIMD each and puts position S + k on the SIMD of about position k
        // (measured).  The older wave of a SIMD takes most of its issue slots,
        // so a SIMD finishes after about the sum of its two tasks, and the
        // slowest SIMD sets the launch time.  In row-major order the second
        // wave is the task 30 strips further south in about the same
        // columns, and costly regions of a mesh (supersonic interfaces run an
        // extra branch; edge windows write ghost cells) are wide and pair up.
        // Reversing the second round pairs each task with its mirror image
        // through the mesh centre instead (2.5% less time per 2000^2 step).
        // The order decides only which wave computes which task.
        if (n == 1 && !newP) mirror_second_round(tasks, 4LL * cu_count);
        const int blocks_needed = (nt + newW - 1) / newW;
        const int newgrid = std::max(1, newP ? std::min(blocks_needed, occ * cu_count) : blocks_needed);
        // Upload into a new buffer; the old task list and descriptors stay
        // valid until every step of this stream that uses them has run.
        HIPCHK(hipStreamSynchronize(stream));
        int4 *dt_new = nullptr;
        HIPCHK(hipMalloc(&dt_new, std::max(1, nt) * sizeof(int4)));
        hipError_t e = hipMemcpyAsync(dt_new, tasks.data(), nt * sizeof(int4), hipMemcpyHostToDevice, stream);
        if (e == hipSuccess)
            e = hipMemcpyAsync(d_desc, nd.data(), n * sizeof(InstDesc), hipMemcpyHostToDevice, stream);
        const hipError_t es = hipStreamSynchronize(stream);
        if (e == hipSuccess) e = es;
        if (e != hipSuccess) {
            hipFree(dt_new);
            // the descriptor copy may have landed partly: restore the current ones
            hipMemcpy(d_desc, hd.data(), n * sizeof(InstDesc), hipMemcpyHostToDevice);
            return fail(FVM2D_ERR_HIP, std::string("uploading the launch layout: ") + hipGetErrorString(e));
        }
        // the captured graph replays the old task list and grid
        drop_graph();
        if (d_tasks) hipFree(d_tasks);
        d_tasks = dt_new;
        hd.swap(nd);
        h_tasks.swap(tasks);
        ntasks = nt;
        R = newR;
        wpb = newW;
        persistent = newP;
        grid = newgrid;
        return FVM2D_OK;
    }

    void drop_graph() {
        if (gexec) {
            hipStreamSynchronize(stream);  // no replay of it may still be queued
            hipGraphExecDestroy(gexec);
        }
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
        // one pass over the vec5d rows, four plane rows written at once
        host_rows(d.ny, [&](int j0, int j1) {
            for (int j = j0; j < j1; j++) {
                const double *s = U5 + (long long)j * d.nx * 5;
                double *dst[4];
                for (int k = 0; k < 4; k++) dst[k] = stage + k * d.plane + (long long)j * d.pitch;
                for (int c = 0; c < d.nx; c++)
                    for (int k = 0; k < 4; k++) dst[k][c] = s[5 * c + k];
                for (int k = 0; k < 4; k++)
                    for (int c = d.nx; c < d.pitch; c++) dst[k][c] = 0.0;
            }
        });
        HIPCHK(hipMemcpyAsync(d.buf0, stage, 4 * d.plane * sizeof(double), hipMemcpyHostToDevice, stream));
        HIPCHK(hipStreamSynchronize(stream));
        begun = false;
        return FVM2D_OK;
    }

    // Buffer holding instance i's current state: buf0 between time loops;
    // inside one (after begin, before finish_run) the step kernel has
    // written buf[steps & 1], steps counted from begin.
    int current_buffer(int i, const double **src) {
        const InstDesc &d = hd[i];
        *src = d.buf0;
        if (begun) {
            int rc = read_ctl();
            if (rc) return rc;
            if (h_ctl[i].steps & 1) *src = d.buf1;
        }
        return FVM2D_OK;
    }

    int download(int i, double *U5) {
        if (i < 0 || i >= n || !U5) return fail(FVM2D_ERR_ARG, "bad instance or null state");
        const double *src = nullptr;
        int rc = current_buffer(i, &src);
        if (rc) return rc;
        return copy_out(i, src, U5);
    }

    // Pitched 4-plane device state of instance i -> host vec5d [ny][nx][5].
    int copy_out(int i, const double *src, double *U5) {
        const InstDesc &d = hd[i];
        HIPCHK(hipMemcpyAsync(stage, src, 4 * d.plane * sizeof(double), hipMemcpyDeviceToHost, stream));
        HIPCHK(hipStreamSynchronize(stream));
        host_rows(d.ny, [&](int j0, int j1) {
            for (int j = j0; j < j1; j++) {
                double *o = U5 + (long long)j * d.nx * 5;
                for (int c = 0; c < d.nx; c++) {
                    for (int k = 0; k < 4; k++) o[5 * c + k] = stage[k * d.plane + (long long)j * d.pitch + c];
                    o[5 * c + 4] = 0.0;
                }
            }
        });
        return FVM2D_OK;
    }

    // fn(j0, j1) over row blocks of [0, ny) on up to 8 host threads (the
    // vec5d <-> plane reordering of large meshes is bound by host memory
    // bandwidth, which one thread does not reach); small meshes stay on the
    // calling thread.
    template <class F>
    static void host_rows(int ny, F &&fn) {
        const unsigned hw = std::max(1u, std::thread::hardware_concurrency());
        const int nt = (int)std::min<long long>({8, (long long)hw, std::max(1, ny / 128)});
        if (nt <= 1) {
            fn(0, ny);
            return;
        }
        auto block = [&fn, nt, ny](int t) {
            fn((int)((long long)ny * t / nt), (int)((long long)ny * (t + 1) / nt));
        };
        std::vector<std::thread> pool;
        int t = 1;
        for (; t < nt; t++) {
            try {
                pool.emplace_back(block, t);
            } catch (...) {  // no thread to be had: the rest runs here
                break;
            }
        }
        for (int r = t; r < nt; r++) block(r);
        block(0);
        for (auto &th : pool) th.join();
    }

    // Record-mode frame of instance i inside a time loop: the state after
    // the last completed step as it was before that step's ghost fill.
    // The step's input is still in the other time-loop buffer (nothing has
    // overwritten it), and dt_used is the step's own dt.
    int frame(int i, double *U5) {
        if (i < 0 || i >= n || !U5) return fail(FVM2D_ERR_ARG, "bad instance or null state");
        const int rc = make_frame(i);
        return rc ? rc : copy_out(i, d_frame, U5);
    }

    // The frame's density as the 8-bit image of the reference's frame file
    // (min-max normalised and rounded on the device), gray[j * nx + i]:
    // nx * ny bytes cross to the host instead of the whole state.
    int frame_gray(int i, unsigned char *gray) {
        if (i < 0 || i >= n || !gray) return fail(FVM2D_ERR_ARG, "bad instance or null image");
        int rc = make_frame(i);
        if (rc) return rc;
        const InstDesc &d = hd[i];
        if (!d_gray) {
            long long mx = 0;
            for (const auto &e : hd) mx = std::max(mx, (long long)e.nx * e.ny);
            HIPCHK(hipMalloc(&d_gray, mx));
            HIPCHK(hipMalloc(&d_range, 2 * sizeof(unsigned long long)));
        }
        launch_frame_gray(d, d_frame, d_range, d_gray, stream);
        HIPCHK(hipGetLastError());
        // through the pinned staging buffer (it holds a whole state)
        const size_t bytes = (size_t)d.nx * d.ny;
        HIPCHK(hipMemcpyAsync(stage, d_gray, bytes, hipMemcpyDeviceToHost, stream));
        HIPCHK(hipStreamSynchronize(stream));
        std::memcpy(gray, stage, bytes);
        return FVM2D_OK;
    }

    // The pre-fill frame of instance i into d_frame (see frame()).
    int make_frame(int i) {
        if (!begun) return fail(FVM2D_ERR_STATE, "a frame needs a time loop (begin/advance)");
        int rc = read_ctl();
        if (rc) return rc;
        const Ctl &c = h_ctl[i];
        if (c.steps < 1) return fail(FVM2D_ERR_STATE, "no step taken yet in this time loop");
        const InstDesc &d = hd[i];
        if (!d_frame) {
            long long mx = 0;
            for (const auto &e : hd) mx = std::max(mx, e.plane);
            HIPCHK(hipMalloc(&d_frame, 4 * mx * sizeof(double)));
        }
        const double *cur = (c.steps & 1) ? d.buf1 : d.buf0;
        const double *prev = (c.steps & 1) ? d.buf0 : d.buf1;
        launch_frame(exact, d, prev, cur, c.dt_used, d_frame, stream);
        HIPCHK(hipGetLastError());
        return FVM2D_OK;
    }

    int planes_device(int i, double *d_planes, bool in) {
        if (i < 0 || i >= n || !d_planes) return fail(FVM2D_ERR_ARG, "bad instance or null pointer");
        const InstDesc &d = hd[i];
        const double *cur = d.buf0;
        if (!in) {
            int rc = current_buffer(i, &cur);
            if (rc) return rc;
        }
        for (int k = 0; k < 4; k++) {
            double *a = (in ? d.buf0 : const_cast<double *>(cur)) + k * d.plane;
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
    // t0: per-instance start times (host array of n values), or null for 0.
    int begin(const double *t0 = nullptr) {
        if (t0) {
            // h_t0 is rewritten only after the stream has consumed it
            HIPCHK(hipStreamSynchronize(stream));
            for (int i = 0; i < n; i++) h_t0[i] = t0[i];
            HIPCHK(hipMemcpyAsync(d_t0, h_t0, n * sizeof(double), hipMemcpyHostToDevice, stream));
        }
        return begin_device(t0 ? d_t0 : nullptr);
    }

    // begin() with the start times in device memory (n values, or null for
    // 0): enqueues the reset without waiting for the host.
    int begin_device(const double *d_times) {
        HIPCHK(hipMemsetAsync(d_ctl, 0, n * sizeof(Ctl), stream));
        for (int i = 0; i < n; i++) launch_speed(exact, d_desc, d_ctl, i, 0, hd[i].ny, stream);
        launch_ctl_start(d_desc, d_ctl, n, d_times, stream);
        HIPCHK(hipGetLastError());
        begun = true;
        return FVM2D_OK;
    }

    // Start time loops from control blocks saved after an earlier loop
    // (device array of n), each instance's state already in buf0.
    int begin_from_ctl(const Ctl *d_saved) {
        launch_ctl_restore(d_ctl, d_saved, n, stream);
        HIPCHK(hipGetLastError());
        begun = true;
        return FVM2D_OK;
    }

    void launch_one() {
        launch_step(exact, d_desc, d_ctl, n, d_tasks, ntasks, d_sched, persistent, grid, wpb, stream);
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

    // Enqueue nsteps step launches: graph replays of kGraphSteps launches
    // (the graph is captured only when at least one replay follows) and a
    // remainder of plain launches.
    int enqueue(int nsteps) {
        int k = 0;
        if (nsteps >= kGraphSteps) {
            int rc = ensure_graph(kGraphSteps);
            if (rc) return rc;
            for (; k + kGraphSteps <= nsteps; k += kGraphSteps) HIPCHK(hipGraphLaunch(gexec, stream));
        }
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

    // Device-resident time loop until every instance is done.  The timed
    // region (ms_out) starts after the initial s_max and dt are known: the
    // first chunk is sized from them (exactly, under a step limit), and a
    // launch that finds its instance done exits at once.
    int run_to_completion(double *ms_out) {
        if (!begun) {
            int rc = begin();
            if (rc) return rc;
        }
        int rc = read_ctl();
        if (rc) return rc;
        int chunk = std::max(1, remaining_estimate());
        // capture the step graph before the timed region when the loop uses it
        if (chunk >= kGraphSteps && (rc = ensure_graph(kGraphSteps))) return rc;
        HIPCHK(hipEventRecord(e0, stream));
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
        // Strip heights: a fixed ladder plus, for k = 1..8 rounds of resident
        // waves, the height that makes the wave count just fit k rounds
        // (a launch whose last round is nearly empty wastes that round).
        std::vector<int> Rs = {4, 6, 8, 10, 12, 16, 20, 24, 32, 48, 64};
        {
            const int occ = std::max(1, step_kernel_occupancy(exact, 4));
            const long long slots = (long long)occ * 4 * cu_count;
            long long cols = 0;  // windows summed over instances, weighted by rows
            long long rows_max = 0;
            for (int i = 0; i < n; i++) {
                cols += hd[i].nw;
                rows_max = std::max<long long>(rows_max, hd[i].ny - 4);
            }
            for (int k = 1; k <= 8 && n == 1; k++) {
                const long long strips = std::max<long long>(1, k * slots / std::max<long long>(1, cols));
                const int r = (int)((rows_max + strips - 1) / strips);
                if (r >= 4 && r <= 96) Rs.push_back(r);
            }
            if (n == 1) {  // half a round: one wave per SIMD
                const long long strips = std::max<long long>(1, slots / 2 / std::max<long long>(1, cols));
                const int r = (int)((rows_max + strips - 1) / strips);
                if (r >= 4 && r <= 96) Rs.push_back(r);
            }
            // several meshes: for k rounds, the lowest height whose task
            // count still fits k rounds (the last round full, not a few
            // waves running alone)
            auto tasks_at = [&](int r) {
                long long t = 0;
                for (int i = 0; i < n; i++) t += (long long)hd[i].nw * ((hd[i].ny - 4 + r - 1) / r);
                return t;
            };
            for (int k = 1; k <= 12 && n > 1; k++) {
                for (int r = 4; r <= 96; r++) {
                    const long long t = tasks_at(r);
                    if (t <= k * slots) {
                        if (t * 10 > k * slots * 8) Rs.push_back(r);
                        break;
                    }
                }
            }
            std::sort(Rs.begin(), Rs.end());
            Rs.erase(std::unique(Rs.begin(), Rs.end()), Rs.end());
        }
        double best = std::numeric_limits<double>::infinity();
        int bR = R, bW = wpb, bP = persistent;
        for (int r : Rs) {
            if (rc) break;
            if (r > std::max(4, maxny)) continue;
            for (int w : {2, 4}) {
                for (int pers : {0, 1}) {
                    rc = set_launch(r, w, pers);
                    if (rc) break;
                    // One mesh in one round of two waves per SIMD: the task
                    // order above pairs mirror-image tasks on a SIMD for the
                    // dispatch of four waves per block; blocks of two waves
                    // land on SIMDs in another pattern.  Trials from the
                    // initial state cannot tell the two apart reliably
                    // (2000 x 2000: blocks of two ahead by 2% over the first
                    // 400 steps, behind by 4% over 1500), so only the
                    // geometry the order was made for is tried.
                    if (n == 1 && pers == 0 && w != 4 && ntasks > 4 * cu_count &&
                        ntasks <= 8 * cu_count)
                        continue;
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
