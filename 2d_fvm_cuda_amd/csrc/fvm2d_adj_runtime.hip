// Adjoint front end of include/fvm2d.h: forward solve with binomial
// time-step checkpointing, then the reverse sweep of every step.
//
// Checkpoint schedule.  cost[n][c] is the least number of forward steps
// needed to reverse n steps whose first state is stored, with c further
// checkpoint slots free:
//     cost[1][c] = 0,  cost[n][0] = n (n - 1) / 2,
//     cost[n][c] = min_{0<m<n} m + cost[n-m][c-1] + cost[m][c],
// the recurrence whose minimum is the binomial (revolve) bound; split[n][c]
// keeps the minimising m.  The forward call stores U_0 and the checkpoints
// on the schedule's first spine; the backward call follows the recursion,
// recomputing from the nearest stored state.
#include <map>

#include "fvm2d_adjoint_kernels.h"
#include "fvm2d_engine.h"

namespace fvm2d {

struct AdjEngine {
    Engine fwd;                      // forward stepping of all instances
    int n = 0, nsteps = 0, slots = 0;
    std::vector<int> nx, ny;
    std::vector<double> gamma;
    // per instance device buffers
    std::vector<double *> ck;        // [inst * slots + s] checkpoint states
    std::vector<double *> A1, B1, lamA1, lamB1, lamX, lamY, UN, U0c;
    double *dtpart = nullptr;
    // state locations (the slots, then the engine's buf0 and buf1) as
    // device tables of n instance pointers each, for one-launch copies
    double **d_locptr = nullptr;
    long long *d_counts = nullptr;
    long long max_count = 0;
    AdjCtl *d_actl = nullptr;
    std::vector<AdjDesc> had;
    int4 *d_tfwd = nullptr, *d_tfy = nullptr, *d_tr2 = nullptr, *d_tr3a = nullptr, *d_tr3b = nullptr;
    int nfwd = 0, nfy = 0, nr2 = 0, nr3a = 0, nr3b = 0;
    // rows per task of the marched B1 = Y(U) recompute
    static constexpr int kFwdYRows = 16;
    std::vector<void *> allocs;
    std::vector<double *> stage_in, stage_out;  // host-pointer API staging
    // schedule
    std::vector<std::vector<long long>> cost;
    std::vector<std::vector<int>> split;
    Ctl *d_slot_ctl = nullptr;       // [slot * n + inst] control block of the stored state
    AdjDesc *d_atab = nullptr;       // reverse-step descriptors, see build_desc_table
    std::vector<int> slot_step;      // step index held by each slot (-1: free)
    std::map<int, int> stored;       // step -> slot
    // stats
    long long recompute_steps = 0;
    long long max_cells = 0;  // largest nx * ny over the instances
    int reverse_steps = 0;
    double fwd_ms = 0.0, bwd_ms = 0.0, rev_ms_sum = 0.0;
    hipEvent_t r0 = nullptr, r1 = nullptr;
    // reverse-step timing: a ring of event pairs read back lazily, so the
    // host never waits for a reverse step before queueing the next one
    static constexpr int kRing = 64;
    hipEvent_t ring0[kRing] = {}, ring1[kRing] = {};
    int ring_pending = 0;  // recorded pairs not yet added to rev_ms_sum
    bool have_forward = false;
    bool launch_chosen = false;  // the forward engine's launch geometry is fixed

    ~AdjEngine() {
        if (fwd.stream) hipStreamSynchronize(fwd.stream);  // reverse steps may still be queued
        for (void *p : allocs) hipFree(p);
        if (r0) hipEventDestroy(r0);
        if (r1) hipEventDestroy(r1);
        for (int k = 0; k < kRing; k++) {
            if (ring0[k]) hipEventDestroy(ring0[k]);
            if (ring1[k]) hipEventDestroy(ring1[k]);
        }
    }

    template <class T>
    int dalloc(T **p, size_t count) {
        HIPCHK(hipMalloc((void **)p, std::max<size_t>(1, count) * sizeof(T)));
        allocs.push_back((void *)*p);
        return FVM2D_OK;
    }

    int init(int count, const int *nx_, const int *ny_, const fvm2d_params *p, int steps, int nslots) {
        if (count < 1 || steps < 1) return fail(FVM2D_ERR_ARG, "need at least one instance and one step");
        n = count;
        nsteps = steps;
        std::vector<fvm2d_params> fp(p, p + count);
        for (auto &q : fp) {
            q.t_end = std::numeric_limits<double>::infinity();
            q.max_steps = -1;
            q.mode = FVM2D_MODE_FAST;
        }
        int rc = fwd.init(count, nx_, ny_, fp.data());
        if (rc) return rc;
        slots = std::max(1, std::min(nslots <= 0 ? steps : nslots, steps));
        nx.assign(nx_, nx_ + count);
        ny.assign(ny_, ny_ + count);
        for (int i = 0; i < count; i++) max_cells = std::max(max_cells, (long long)nx[i] * ny[i]);
        HIPCHK(hipEventCreate(&r0));
        HIPCHK(hipEventCreate(&r1));
        for (int k = 0; k < kRing; k++) {
            HIPCHK(hipEventCreate(&ring0[k]));
            HIPCHK(hipEventCreate(&ring1[k]));
        }
        had.resize(n);
        ck.resize((size_t)n * slots);
        A1.resize(n), B1.resize(n), lamA1.resize(n), lamB1.resize(n), lamX.resize(n), lamY.resize(n);
        UN.resize(n), U0c.resize(n);
        rc = dalloc(&d_actl, n);
        if (rc) return rc;
        for (int i = 0; i < n; i++) {
            const InstDesc &d = fwd.hd[i];
            const size_t st = 4 * (size_t)d.plane;
            for (int s = 0; s < slots; s++)
                if ((rc = dalloc(&ck[(size_t)i * slots + s], st))) return rc;
            for (double **b : {&A1[i], &B1[i], &lamA1[i], &lamB1[i], &lamX[i], &lamY[i], &UN[i]})
                if ((rc = dalloc(b, st))) return rc;
            if ((rc = dalloc(&U0c[i], 4 * (size_t)nx[i] * ny[i]))) return rc;
            HIPCHK(hipMemsetAsync(A1[i], 0, st * sizeof(double), fwd.stream));
            HIPCHK(hipMemsetAsync(B1[i], 0, st * sizeof(double), fwd.stream));
            AdjDesc &a = had[i];
            a.nx = d.nx;
            a.ny = d.ny;
            a.pitch = d.pitch;
            a.strategy = d.strategy;
            a.plane = d.plane;
            a.cfl = d.cfl;
            a.hmin = d.hmin;
            a.lx = d.lx;
            a.ly = d.ly;
            a.gas = d.gas;
            a.A1 = A1[i];
            a.B1 = B1[i];
            a.lamA1 = lamA1[i];
            a.lamB1 = lamB1[i];
            a.ctl = d_actl + i;
            gamma.push_back(p[i].gamma);
        }
        rc = build_tasks();
        if (rc) return rc;
        rc = dalloc(&d_slot_ctl, (size_t)slots * n);
        if (!rc) rc = dalloc(&d_atab, (size_t)(slots + 2) * 2 * n);
        if (!rc) rc = dalloc(&d_locptr, (size_t)(slots + 2) * n);
        if (!rc) rc = dalloc(&d_counts, (size_t)n);
        if (rc) return rc;
        {
            std::vector<double *> lp((size_t)(slots + 2) * n);
            std::vector<long long> cnt(n);
            for (int loc = 0; loc < slots + 2; loc++)
                for (int i = 0; i < n; i++)
                    lp[(size_t)loc * n + i] = const_cast<double *>(state_ptr(i, loc_where(loc)));
            for (int i = 0; i < n; i++) {
                cnt[i] = 4 * fwd.hd[i].plane;
                max_count = std::max(max_count, cnt[i]);
            }
            HIPCHK(hipMemcpy(d_locptr, lp.data(), lp.size() * sizeof(double *), hipMemcpyHostToDevice));
            HIPCHK(hipMemcpy(d_counts, cnt.data(), cnt.size() * sizeof(long long), hipMemcpyHostToDevice));
        }
        build_schedule();
        HIPCHK(hipStreamSynchronize(fwd.stream));
        return FVM2D_OK;
    }

    int build_tasks() {
        std::vector<int4> tf, tfy, t2, t3a, t3b;
        // B1 = Y(U) by the register march (strips of kFwdYRows interior
        // rows of 64 whole columns: row-segment loads) when its strips fill
        // two waves per SIMD; otherwise by column lines like A1 (more, smaller
        // tasks: at 512 x 512 the march has 256 strips and its reverse step
        // took 5 % longer; on the 32-mesh batch it saves 7 % of the backward)
        long long strips = 0;
        for (int i = 0; i < n; i++)
            strips += (long long)((nx[i] + kWave - 1) / kWave) *
                      ((ny[i] - 4 + kFwdYRows - 1) / kFwdYRows);
        const bool march = strips >= 2LL * 4 * fwd.cu_count;
        for (int i = 0; i < n; i++) {
            const int wf_x = (nx[i] + 59) / 60, wf_y = (ny[i] + 59) / 60;
            for (int j = 0; j < ny[i]; j++)
                for (int w = 0; w < wf_x; w++) tf.push_back(make_int4(i, 0, w, j));
            if (march) {
                const int H = ny[i] - 4, wy = (nx[i] + kWave - 1) / kWave;
                for (int r0 = 2; r0 < 2 + H; r0 += kFwdYRows)
                    for (int w = 0; w < wy; w++)
                        tfy.push_back(make_int4(i, w, r0, std::min(2 + H, r0 + kFwdYRows)));
            } else {
                // Y lines: adjacent columns consecutive, so the waves of a
                // block (and of neighbouring blocks) share each row's lines
                for (int w = 0; w < wf_y; w++)
                    for (int c = 0; c < nx[i]; c++) tf.push_back(make_int4(i, 1, w, c));
            }
        }
        int base = 0;
        for (int i = 0; i < n; i++) {  // R2: X^T on B1 (rows) and Y^T on A1 (columns)
            had[i].part_lo[0] = base + (int)t2.size();
            const int wa_x = (nx[i] + 55) / 56, wa_y = (ny[i] + 55) / 56;
            for (int j = 0; j < ny[i]; j++)
                for (int w = 0; w < wa_x; w++) t2.push_back(make_int4(i, 0, w, j));
            for (int w = 0; w < wa_y; w++)
                for (int c = 0; c < nx[i]; c++) t2.push_back(make_int4(i, 1, w, c));
            had[i].part_hi[0] = base + (int)t2.size();
        }
        base += (int)t2.size();
        const int base3a = base;
        for (int i = 0; i < n; i++) {  // R3a: X^T on U (rows)
            had[i].part_lo[1] = base3a + (int)t3a.size();
            const int wa_x = (nx[i] + 55) / 56;
            for (int j = 0; j < ny[i]; j++)
                for (int w = 0; w < wa_x; w++) t3a.push_back(make_int4(i, 0, w, j));
            had[i].part_hi[1] = base3a + (int)t3a.size();
        }
        base += (int)t3a.size();
        const int base3b = base;
        for (int i = 0; i < n; i++) {  // R3b: Y^T on U (columns)
            had[i].part_lo[2] = base3b + (int)t3b.size();
            const int wa_y = (ny[i] + 55) / 56;
            for (int w = 0; w < wa_y; w++)
                for (int c = 0; c < nx[i]; c++) t3b.push_back(make_int4(i, 1, w, c));
            had[i].part_hi[2] = base3b + (int)t3b.size();
        }
        base += (int)t3b.size();
        nfwd = (int)tf.size();
        nfy = (int)tfy.size();
        nr2 = (int)t2.size();
        nr3a = (int)t3a.size();
        nr3b = (int)t3b.size();
        int rc;
        if ((rc = dalloc(&dtpart, base))) return rc;
        if ((rc = dalloc(&d_tfwd, nfwd))) return rc;
        if ((rc = dalloc(&d_tfy, nfy))) return rc;
        if ((rc = dalloc(&d_tr2, nr2))) return rc;
        if ((rc = dalloc(&d_tr3a, nr3a))) return rc;
        if ((rc = dalloc(&d_tr3b, nr3b))) return rc;
        for (int i = 0; i < n; i++) had[i].dtpart = dtpart;
        // the host vectors are locals: wait for the copies on every path
        hipError_t e = hipMemcpyAsync(d_tfwd, tf.data(), nfwd * sizeof(int4), hipMemcpyHostToDevice,
                                      fwd.stream);
        if (e == hipSuccess)
            e = hipMemcpyAsync(d_tfy, tfy.data(), nfy * sizeof(int4), hipMemcpyHostToDevice, fwd.stream);
        if (e == hipSuccess)
            e = hipMemcpyAsync(d_tr2, t2.data(), nr2 * sizeof(int4), hipMemcpyHostToDevice, fwd.stream);
        if (e == hipSuccess)
            e = hipMemcpyAsync(d_tr3a, t3a.data(), nr3a * sizeof(int4), hipMemcpyHostToDevice,
                               fwd.stream);
        if (e == hipSuccess)
            e = hipMemcpyAsync(d_tr3b, t3b.data(), nr3b * sizeof(int4), hipMemcpyHostToDevice,
                               fwd.stream);
        const hipError_t es = hipStreamSynchronize(fwd.stream);
        if (e == hipSuccess) e = es;
        if (e != hipSuccess)
            return fail(FVM2D_ERR_HIP, std::string("uploading the adjoint task lists: ") +
                                           hipGetErrorString(e));
        return FVM2D_OK;
    }

    void build_schedule() {
        const int N = nsteps, C = slots - 1;  // slot 0 always holds U_0
        cost.assign(N + 1, std::vector<long long>(C + 1, 0));
        split.assign(N + 1, std::vector<int>(C + 1, 0));
        for (int m = 1; m <= N; m++) cost[m][0] = (long long)m * (m - 1) / 2;
        for (int c = 1; c <= C; c++) {
            cost[1][c] = 0;
            for (int m = 2; m <= N; m++) {
                long long best = -1;
                int bs = 1;
                for (int s = 1; s < m; s++) {
                    const long long v = s + cost[m - s][c - 1] + cost[s][c];
                    if (best < 0 || v < best) {
                        best = v;
                        bs = s;
                    }
                }
                cost[m][c] = best;
                split[m][c] = bs;
            }
        }
    }

    // State locations: a slot index s >= 0, or one of the forward engine's
    // two time-loop buffers, kBuf0 (buf0) and kBuf1 (buf1).
    static constexpr int kBuf0 = -1, kBuf1 = -2;

    const double *state_ptr(int i, int where) const {
        const InstDesc &d = fwd.hd[i];
        return where == kBuf0 ? d.buf0 : where == kBuf1 ? d.buf1 : ck[(size_t)i * slots + where];
    }

    // Location index in the pointer tables: slots first, then buf0, buf1.
    int loc_index(int where) const { return where >= 0 ? where : (where == kBuf0 ? slots : slots + 1); }
    static int loc_where_of(int loc, int nslots) {
        return loc < nslots ? loc : (loc == nslots ? kBuf0 : kBuf1);
    }
    int loc_where(int loc) const { return loc_where_of(loc, slots); }

    // Copy the state of every instance from src to dst (locations as
    // above), all instances in one launch.
    int copy_state(int src, int dst) {
        adj_launch_copy(d_locptr + (size_t)loc_index(src) * n, d_locptr + (size_t)loc_index(dst) * n,
                        d_counts, n, max_count, fwd.stream);
        HIPCHK(hipGetLastError());
        return FVM2D_OK;
    }

    // Advance every instance by m >= 1 steps from the state in slot src,
    // whose control blocks are d_slot_ctl[src]; *res receives the engine buffer that
    // holds the result (kBuf0 or kBuf1: the time loop alternates them) and
    // the engine's control blocks its times.  Everything is queued on the
    // stream: the host does not wait.
    int advance(int src, int m, int *res) {
        int rc = copy_state(src, kBuf0);
        if (!rc) rc = fwd.begin_from_ctl(d_slot_ctl + (size_t)src * n);
        if (!rc) rc = fwd.enqueue(m);
        if (rc) return rc;
        fwd.begun = false;
        recompute_steps += m;
        *res = (m & 1) ? kBuf1 : kBuf0;
        return FVM2D_OK;
    }

    // Store the state at location src into slot s together with the
    // engine's control blocks after the advance that produced it.
    int store_slot(int src, int s) {
        int rc = copy_state(src, s);
        if (rc) return rc;
        HIPCHK(hipMemcpyAsync(d_slot_ctl + (size_t)s * n, fwd.d_ctl, n * sizeof(Ctl),
                              hipMemcpyDeviceToDevice, fwd.stream));
        return FVM2D_OK;
    }

    // Descriptors of every reverse step, built once per backward call: for
    // each state location (the slots, then buf0 and buf1) and each parity of
    // the adjoint buffers (lam_in = lamX, lam_out = lamY, then swapped), the
    // n instance descriptors.
    int build_desc_table() {
        std::vector<AdjDesc> tab((size_t)(slots + 2) * 2 * n);
        for (int loc = 0; loc < slots + 2; loc++)
            for (int par = 0; par < 2; par++)
                for (int i = 0; i < n; i++) {
                    AdjDesc a = had[i];
                    a.U = state_ptr(i, loc_where(loc));
                    // a stored state's control block was saved with it; a
                    // recomputed one is the engine's current block (a
                    // reverse step follows the advance that produced it)
                    a.fctl = loc < slots ? d_slot_ctl + (size_t)loc * n + i : fwd.d_ctl + i;
                    a.lam_in = par ? lamY[i] : lamX[i];
                    a.lam_out = par ? lamX[i] : lamY[i];
                    tab[((size_t)loc * 2 + par) * n + i] = a;
                }
        HIPCHK(hipMemcpy(d_atab, tab.data(), tab.size() * sizeof(AdjDesc), hipMemcpyHostToDevice));
        return FVM2D_OK;
    }

    // Add the elapsed times of the oldest pending reverse steps to
    // rev_ms_sum until at most keep remain pending.
    int drain_ring(int keep) {
        while (ring_pending > keep) {
            const int k = (reverse_steps - ring_pending) % kRing;
            HIPCHK(hipEventSynchronize(ring1[k]));
            float ms = 0.f;
            HIPCHK(hipEventElapsedTime(&ms, ring0[k], ring1[k]));
            rev_ms_sum += ms;
            ring_pending--;
        }
        return FVM2D_OK;
    }

    int take_slot() const {
        for (int s = 1; s < slots; s++)
            if (slot_step[s] < 0) return s;
        return -1;
    }

    // Reverse step n (U_n at location s): lamX (dL/dU_{n+1}) -> lamY
    // (dL/dU_n); then the two are swapped.  The descriptors come from the
    // table (the parity of the reverse steps done so far selects the
    // adjoint buffers, which started as lamX -> lamY).
    int reverse_step(int s) {
        int rc = drain_ring(kRing - 1);
        if (rc) return rc;
        const int k = reverse_steps % kRing;
        const int loc = s >= 0 ? s : (s == kBuf0 ? slots : slots + 1);
        AdjDesc *descs = d_atab + ((size_t)loc * 2 + (reverse_steps & 1)) * n;
        HIPCHK(hipEventRecord(ring0[k], fwd.stream));
        // two resident waves per SIMD for the sweep adjoints (256
        // registers each), four for the lighter forward recompute; the
        // column sweep of R3 writes lam_out and the row sweep adds to it
        const int simds = 4 * fwd.cu_count;
        adj_launch_fwd(descs, d_tfwd, nfwd, 4 * simds, fwd.stream);
        adj_launch_fwd_y(descs, d_tfy, nfy, fwd.stream);
        adj_launch_sweep(descs, d_tr2, nr2, 0, 0, 2 * simds, fwd.stream);
        adj_launch_sweep(descs, d_tr3b, nr3b, 2, nr2 + nr3a, 2 * simds, fwd.stream);
        adj_launch_sweep(descs, d_tr3a, nr3a, 1, nr2, 2 * simds, fwd.stream);
        adj_launch_dt_term(descs, n, nr2 + nr3a + nr3b, fwd.stream);
        HIPCHK(hipGetLastError());
        HIPCHK(hipEventRecord(ring1[k], fwd.stream));
        ring_pending++;
        reverse_steps++;
        std::swap(lamX, lamY);
        return FVM2D_OK;
    }

    // Reverse steps a .. a+len-1; the state U_a is in slot sa.
    int reverse(int a, int len, int c, int sa) {
        if (len == 1) return reverse_step(sa);
        if (c > 0) {
            const int m = split[len][std::min(c, (int)split[len].size() - 1)];
            int s = -1;
            auto it = stored.find(a + m);
            if (it != stored.end()) {
                s = it->second;  // placed by the forward call's spine
            } else if ((s = take_slot()) >= 0) {
                int res;
                int rc = advance(sa, m, &res);
                if (!rc) rc = store_slot(res, s);
                if (rc) return rc;
                slot_step[s] = a + m;
                stored[a + m] = s;
            }
            if (s >= 0) {
                int rc = reverse(a + m, len - m, c - 1, s);
                if (rc) return rc;
                stored.erase(a + m);
                slot_step[s] = -1;
                return reverse(a, m, c, sa);
            }
        }
        // no checkpoint to spare: recompute every state from U_a
        for (int k = len - 1; k >= 1; k--) {
            int res;
            int rc = advance(sa, k, &res);
            if (!rc) rc = reverse_step(res);
            if (rc) return rc;
        }
        return reverse_step(sa);
    }

    int forward(const double *const *U0_cons, double *const *UN_cons) {
        for (int i = 0; i < n; i++) {
            if (!U0_cons[i] || !UN_cons[i]) return fail(FVM2D_ERR_ARG, "null state pointer");
            const InstDesc &d = fwd.hd[i];
            HIPCHK(hipMemcpyAsync(U0c[i], U0_cons[i], 4 * (size_t)nx[i] * ny[i] * sizeof(double),
                                  hipMemcpyDeviceToDevice, fwd.stream));
            adj_launch_cons_to_prim(U0c[i], ck[(size_t)i * slots], nx[i], ny[i], d.pitch, d.plane,
                                    gamma[i] - 1.0, fwd.stream);
        }
        HIPCHK(hipGetLastError());
        if (!launch_chosen) {
            // first forward call: time the candidate launch geometries of the
            // step kernel on this initial state (the autotuner restores it)
            int rc = copy_state(0, kBuf0);
            if (!rc) rc = fwd.autotune(2, nullptr);
            if (rc) return rc;
            launch_chosen = true;
        }
        slot_step.assign(slots, -1);
        stored.clear();
        slot_step[0] = 0;
        stored[0] = 0;
        recompute_steps = 0;
        HIPCHK(hipEventRecord(r0, fwd.stream));
        {
            // control blocks of U_0 (t = 0, s_max and dt of the initial
            // state) for the recomputations that start from slot 0
            int rc = copy_state(0, kBuf0);
            if (!rc) rc = fwd.begin_device(nullptr);
            if (rc) return rc;
            fwd.begun = false;
            HIPCHK(hipMemcpyAsync(d_slot_ctl, fwd.d_ctl, n * sizeof(Ctl), hipMemcpyDeviceToDevice,
                                  fwd.stream));
        }
        // walk the first spine of the schedule, storing its checkpoints
        int a = 0, len = nsteps, c = slots - 1, sa = 0;
        while (len > 1 && c > 0) {
            const int m = split[len][c];
            const int s = take_slot();
            int res;
            int rc = advance(sa, m, &res);
            if (!rc) rc = store_slot(res, s);
            if (rc) return rc;
            slot_step[s] = a + m;
            stored[a + m] = s;
            a += m;
            len -= m;
            c -= 1;
            sa = s;
        }
        int res;
        int rc = advance(sa, len, &res);
        if (rc) return rc;
        HIPCHK(hipEventRecord(r1, fwd.stream));
        rc = fwd.read_ctl();
        if (rc) return rc;
        for (int i = 0; i < n; i++)
            if (fwd.h_ctl[i].status & 1) return fail(FVM2D_ERR_NONFINITE, "non-finite state in the forward run");
        float ms = 0.f;
        HIPCHK(hipEventElapsedTime(&ms, r0, r1));
        fwd_ms = ms;
        for (int i = 0; i < n; i++) {
            const InstDesc &d = fwd.hd[i];
            const double *uN = state_ptr(i, res);
            HIPCHK(hipMemcpyAsync(UN[i], uN, 4 * d.plane * sizeof(double), hipMemcpyDeviceToDevice,
                                  fwd.stream));
            adj_launch_prim_to_cons(uN, UN_cons[i], nx[i], ny[i], d.pitch, d.plane, 1.0 / (gamma[i] - 1.0),
                                    fwd.stream);
        }
        HIPCHK(hipGetLastError());
        HIPCHK(hipStreamSynchronize(fwd.stream));
        forward_recompute_base = recompute_steps;
        have_forward = true;
        return FVM2D_OK;
    }
    long long forward_recompute_base = 0;

    int backward(const double *const *gN_cons, double *const *g0_cons) {
        if (!have_forward) return fail(FVM2D_ERR_STATE, "backward called before forward");
        for (int i = 0; i < n; i++) {
            if (!gN_cons[i] || !g0_cons[i]) return fail(FVM2D_ERR_ARG, "null gradient pointer");
            const InstDesc &d = fwd.hd[i];
            HIPCHK(hipMemsetAsync(lamX[i], 0, 4 * d.plane * sizeof(double), fwd.stream));
            adj_launch_cons_grad_to_prim(UN[i], gN_cons[i], lamX[i], nx[i], ny[i], d.pitch, d.plane,
                                         1.0 / (gamma[i] - 1.0), fwd.stream);
        }
        HIPCHK(hipGetLastError());
        recompute_steps = 0;
        reverse_steps = 0;
        rev_ms_sum = 0.0;
        ring_pending = 0;
        int rc = build_desc_table();
        if (rc) return rc;
        {
            std::vector<AdjCtl> init(n, AdjCtl{0.0, kNoCell, 0, 0});
            HIPCHK(hipMemcpy(d_actl, init.data(), n * sizeof(AdjCtl), hipMemcpyHostToDevice));
        }
        hipEvent_t b0, b1;
        HIPCHK(hipEventCreate(&b0));
        HIPCHK(hipEventCreate(&b1));
        HIPCHK(hipEventRecord(b0, fwd.stream));
        rc = reverse(0, nsteps, slots - 1, 0);
        if (!rc) rc = drain_ring(0);
        if (rc) {
            hipEventDestroy(b0);
            hipEventDestroy(b1);
            return rc;
        }
        for (int i = 0; i < n; i++) {
            const InstDesc &d = fwd.hd[i];
            adj_launch_prim_grad_to_cons(U0c[i], lamX[i], g0_cons[i], nx[i], ny[i], d.pitch, d.plane,
                                         gamma[i] - 1.0, fwd.stream);
        }
        HIPCHK(hipGetLastError());
        HIPCHK(hipEventRecord(b1, fwd.stream));
        HIPCHK(hipEventSynchronize(b1));
        float ms = 0.f;
        HIPCHK(hipEventElapsedTime(&ms, b0, b1));
        bwd_ms = ms;
        hipEventDestroy(b0);
        hipEventDestroy(b1);
        // every forward state used by the schedule after the forward call
        std::vector<AdjCtl> ac(n);
        HIPCHK(hipMemcpy(ac.data(), d_actl, n * sizeof(AdjCtl), hipMemcpyDeviceToHost));
        for (int i = 0; i < n; i++)
            if (ac[i].status & 1)
                return fail(FVM2D_ERR_NONFINITE,
                            "non-finite wave speed in the reverse sweep, or no cell attaining s_max");
        return FVM2D_OK;
    }
};

}  // namespace fvm2d

using namespace fvm2d;

struct fvm2d_adj {
    AdjEngine eng;
};

extern "C" {

int fvm2d_adj_create(int n, const int *nx, const int *ny, const fvm2d_params *p, int nsteps,
                     int checkpoint_slots, fvm2d_adj **out) {
    if (!out || !nx || !ny || !p) return fail(FVM2D_ERR_ARG, "null argument");
    *out = nullptr;
    auto *a = new fvm2d_adj();
    int rc = a->eng.init(n, nx, ny, p, nsteps, checkpoint_slots);
    if (rc) {
        delete a;
        return rc;
    }
    *out = a;
    return FVM2D_OK;
}

int fvm2d_adj_forward(fvm2d_adj *a, const double *const *d_U0_cons, double *const *d_UN_cons) {
    if (!a || !d_U0_cons || !d_UN_cons) return fail(FVM2D_ERR_ARG, "null argument");
    return a->eng.forward(d_U0_cons, d_UN_cons);
}

int fvm2d_adj_backward(fvm2d_adj *a, const double *const *d_gN_cons, double *const *d_g0_cons) {
    if (!a || !d_gN_cons || !d_g0_cons) return fail(FVM2D_ERR_ARG, "null argument");
    return a->eng.backward(d_gN_cons, d_g0_cons);
}

// Host-pointer variants: stage through per-instance device buffers.
static int adj_host_call(fvm2d_adj *a, const double *const *in, double *const *out, bool fwd) {
    if (!a || !in || !out) return fail(FVM2D_ERR_ARG, "null argument");
    AdjEngine &E = a->eng;
    if (E.stage_in.empty()) {
        E.stage_in.resize(E.n);
        E.stage_out.resize(E.n);
        for (int i = 0; i < E.n; i++) {
            const size_t cnt = 4 * (size_t)E.nx[i] * E.ny[i];
            int rc = E.dalloc(&E.stage_in[i], cnt);
            if (!rc) rc = E.dalloc(&E.stage_out[i], cnt);
            if (rc) return rc;
        }
    }
    for (int i = 0; i < E.n; i++) {
        if (!in[i] || !out[i]) return fail(FVM2D_ERR_ARG, "null host pointer");
        HIPCHK(hipMemcpy(E.stage_in[i], in[i], 4 * (size_t)E.nx[i] * E.ny[i] * sizeof(double),
                         hipMemcpyHostToDevice));
    }
    const int rc = fwd ? E.forward(E.stage_in.data(), E.stage_out.data())
                       : E.backward(E.stage_in.data(), E.stage_out.data());
    if (rc) return rc;
    for (int i = 0; i < E.n; i++)
        HIPCHK(hipMemcpy(out[i], E.stage_out[i], 4 * (size_t)E.nx[i] * E.ny[i] * sizeof(double),
                         hipMemcpyDeviceToHost));
    return FVM2D_OK;
}

int fvm2d_adj_forward_host(fvm2d_adj *a, const double *const *U0_cons, double *const *UN_cons) {
    return adj_host_call(a, U0_cons, UN_cons, true);
}

int fvm2d_adj_backward_host(fvm2d_adj *a, const double *const *gN_cons, double *const *g0_cons) {
    return adj_host_call(a, gN_cons, g0_cons, false);
}

int fvm2d_adj_get_stats(const fvm2d_adj *a, fvm2d_adj_stats *st) {
    if (!a || !st) return fail(FVM2D_ERR_ARG, "null argument");
    const AdjEngine &E = a->eng;
    st->forward_ms = E.fwd_ms;
    st->backward_ms = E.bwd_ms;
    st->reverse_step_ms = E.reverse_steps ? E.rev_ms_sum / E.reverse_steps : 0.0;
    st->forward_step_ms = E.nsteps ? E.fwd_ms / E.nsteps : 0.0;
    st->reverse_steps = E.reverse_steps;
    st->recompute_steps = (int)E.recompute_steps;
    st->checkpoint_slots = E.slots;
    st->schedule_cost = (int)E.cost[E.nsteps][E.slots - 1];
    return FVM2D_OK;
}

int fvm2d_adj_set_launch(fvm2d_adj *a, int strip_rows, int waves_per_block, int persistent) {
    if (!a) return fail(FVM2D_ERR_ARG, "adjoint is null");
    a->eng.launch_chosen = true;
    return a->eng.fwd.set_launch(strip_rows, waves_per_block, persistent);
}

int fvm2d_adj_destroy(fvm2d_adj *a) {
    delete a;
    return FVM2D_OK;
}

}  // extern "C"
