// C ABI of include/fvm2d.h for the plain and batched forward solvers.
#include <cmath>
#include <vector>

#include "fvm2d_engine.h"

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

int fvm2d_download_frame(fvm2d_solver *s, double *U5) {
    if (!s) return fail(FVM2D_ERR_ARG, "solver is null");
    return s->eng.frame(0, U5);
}

int fvm2d_download_frame_gray(fvm2d_solver *s, unsigned char *gray) {
    if (!s) return fail(FVM2D_ERR_ARG, "solver is null");
    return s->eng.frame_gray(0, gray);
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
    // capture the graph (if the steps use it) outside the timed region
    int rc = nsteps >= Engine::kGraphSteps ? E.ensure_graph(Engine::kGraphSteps) : FVM2D_OK;
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

int fvm2d_time_kernel(fvm2d_solver *s, int nsteps, double *avg_ms_out, double *min_ms_out) {
    if (!s) return fail(FVM2D_ERR_ARG, "solver is null");
    if (nsteps < 1) return fail(FVM2D_ERR_ARG, "nsteps must be >= 1");
    Engine &E = s->eng;
    if (!E.begun) {
        int rc = E.begin();
        if (rc) return rc;
    }
    // average: nsteps back-to-back launches between two events
    hipEvent_t a, b;
    HIPCHK(hipEventCreate(&a));
    HIPCHK(hipEventCreate(&b));
    HIPCHK(hipEventRecord(a, E.stream));
    for (int k = 0; k < nsteps; k++) E.launch_one();
    HIPCHK(hipEventRecord(b, E.stream));
    HIPCHK(hipEventSynchronize(b));
    float total = 0.f;
    HIPCHK(hipEventElapsedTime(&total, a, b));
    hipEventDestroy(a);
    hipEventDestroy(b);
    // minimum: launches bracketed one by one
    std::vector<hipEvent_t> ev(2 * nsteps);
    for (auto &e : ev) HIPCHK(hipEventCreate(&e));
    for (int k = 0; k < nsteps; k++) {
        HIPCHK(hipEventRecord(ev[2 * k], E.stream));
        E.launch_one();
        HIPCHK(hipEventRecord(ev[2 * k + 1], E.stream));
    }
    HIPCHK(hipEventSynchronize(ev[2 * nsteps - 1]));
    double mn = 1e300;
    for (int k = 0; k < nsteps; k++) {
        float ms = 0.f;
        HIPCHK(hipEventElapsedTime(&ms, ev[2 * k], ev[2 * k + 1]));
        mn = std::min(mn, (double)ms);
    }
    for (auto &e : ev) hipEventDestroy(e);
    if (avg_ms_out) *avg_ms_out = total / nsteps;
    if (min_ms_out) *min_ms_out = mn;
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
    int rc = nsteps >= Engine::kGraphSteps ? E.ensure_graph(Engine::kGraphSteps) : FVM2D_OK;
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

int fvm2d_strip_layout(int nx, int ny, int strip_rows, int waves_per_block, int persistent,
                       int cu_count, int blocks_per_cu, int *out, int max_tasks, int *ntasks_out) {
    if (nx < 6 || ny < 6 || strip_rows < 1 || waves_per_block < 1 || waves_per_block > 4 ||
        cu_count < 1 || blocks_per_cu < 1 || max_tasks < 0 || (max_tasks > 0 && !out) ||
        !ntasks_out)
        return fail(FVM2D_ERR_ARG, "bad layout query");
    const int nw = (nx - 4 + kOwn - 1) / kOwn;
    if (2LL * nw * ((ny - 4 + strip_rows - 1) / strip_rows) > (1LL << 30))
        return fail(FVM2D_ERR_ARG, "strip rows too low for this mesh");
    const long long slots = (long long)blocks_per_cu * waves_per_block * cu_count;
    std::vector<int4> tasks;
    Engine::mesh_strips(0, nw, ny, strip_rows, persistent == 0, slots, 4LL * cu_count,
                        waves_per_block, tasks);
    if (persistent == 0) Engine::mirror_second_round(tasks, 4LL * cu_count);
    *ntasks_out = (int)tasks.size();
    for (int k = 0; k < (int)tasks.size() && k < max_tasks; k++) {
        out[3 * k] = tasks[k].y;
        out[3 * k + 1] = tasks[k].z;
        out[3 * k + 2] = tasks[k].w;
    }
    return FVM2D_OK;
}

}  // extern "C"
