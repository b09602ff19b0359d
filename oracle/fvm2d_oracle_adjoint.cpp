// FP64 host reference gradient of the split-step solve.  TEST INFRASTRUCTURE
// ONLY (the checker of the GPU adjoint, never the product).
//
// The function differentiated is the solve written with the FAST per-cell
// numerics of the product (2d_fvm_cuda_amd/csrc/fvm2d_cell.h), whose
// per-interface derivatives (fvm2d_adjoint_cell.h) are pinned separately
// against finite differences on the host.  Everything else here is an
// independent plain-loop restatement: whole-line sweeps and their adjoints,
// the time loop of src/wrappers.cuh:196-284 with every state stored, the
// ghost fill of src/ghostCellPresets.cuh:47-241 run as its two phases and
// reversed phase by phase, and the dt = CFL h / s_max dependence through
// the first cell attaining s_max.
#include <cmath>
#include <cstring>
#include <vector>

#include "../2d_fvm_cuda_amd/csrc/fvm2d_adjoint_cell.h"
#include "fvm2d_oracle.h"

using namespace fvm2d;

namespace {

struct Mesh {
    int nx, ny;
    Line lx, ly;
    Gas gas;
    double cfl, hmin;
    int strategy;
};

// Q at (j, i) in X order
using Field = std::vector<Q>;

inline Q &at(Field &f, const Mesh &m, int j, int i) { return f[(size_t)j * m.nx + i]; }
inline const Q &at(const Field &f, const Mesh &m, int j, int i) { return f[(size_t)j * m.nx + i]; }
inline Q swapq(const Q &q) { return Q{q.r, q.p, q.ut, q.un}; }

// Line sweep: cells k in [2, n-3] updated, others copied.
void line_fwd(const std::vector<Q> &q, std::vector<Q> &out, const Line &L, const Gas &G, double dt) {
    const int n = (int)q.size();
    std::vector<Q> s(n, Q{0, 0, 0, 0}), fm(n), fp(n), F(n + 1, Q{0, 0, 0, 0});
    for (int k = 1; k <= n - 2; k++) {
        s[k] = slope<false>(q[k - 1], q[k], q[k + 1], L);
        fm[k] = face_minus<false>(q[k], s[k], L);
        fp[k] = face_plus<false>(q[k], s[k], L);
    }
    for (int i = 2; i <= n - 2; i++) {
        IfState W, Wt;
        grp<false>(fp[i - 1], fm[i], s[i - 1], s[i], G, W, Wt);
        F[i] = flux<false>(W, Wt, dt);
    }
    out = q;
    const double lam = dt / L.h;
    for (int k = 2; k <= n - 3; k++) out[k] = update<false>(q[k], F[k], F[k + 1], lam, G);
}

// Adjoint of line_fwd: qb = (d out / d q)^T qbo; adds d(out . qbo)/d dt to dtb.
void line_adj(const std::vector<Q> &q, const std::vector<Q> &qbo, std::vector<Q> &qb, const Line &L,
              const Gas &G, double dt, double &dtb) {
    const int n = (int)q.size();
    const Q Z{0, 0, 0, 0};
    std::vector<Q> s(n, Z), fm(n, Z), fp(n, Z), F(n + 1, Z);
    std::vector<IfState> W(n + 1), Wt(n + 1);
    for (int k = 1; k <= n - 2; k++) {
        s[k] = slope<false>(q[k - 1], q[k], q[k + 1], L);
        fm[k] = face_minus<false>(q[k], s[k], L);
        fp[k] = face_plus<false>(q[k], s[k], L);
    }
    for (int i = 2; i <= n - 2; i++) {
        grp<false>(fp[i - 1], fm[i], s[i - 1], s[i], G, W[i], Wt[i]);
        F[i] = flux<false>(W[i], Wt[i], dt);
    }
    const double lam = dt / L.h;
    std::vector<Q> Ub(n, Z), fmb(n, Z), fpb(n, Z), sb(n, Z);
    qb.assign(n, Z);
    for (int k = 0; k < n; k++) {
        if (k >= 2 && k <= n - 3) {
            const Q qo = update<false>(q[k], F[k], F[k + 1], lam, G);
            Ub[k] = prim_adj(qo, qbo[k], G);
            dtb -= ((F[k + 1].r - F[k].r) * Ub[k].r + (F[k + 1].p - F[k].p) * Ub[k].p +
                    (F[k + 1].un - F[k].un) * Ub[k].un + (F[k + 1].ut - F[k].ut) * Ub[k].ut) *
                   L.ih;
            const Q c = cons_adj(q[k], Ub[k], G);
            qb[k] = c;
        } else {
            qb[k] = qbo[k];
        }
    }
    for (int i = 2; i <= n - 2; i++) {
        // F_i enters cell i (as the left flux, +lam) and cell i-1 (right flux, -lam)
        Q Fb{lam * (Ub[i].r - Ub[i - 1].r), lam * (Ub[i].p - Ub[i - 1].p), lam * (Ub[i].un - Ub[i - 1].un),
             lam * (Ub[i].ut - Ub[i - 1].ut)};
        IfState Wb, Wtb;
        dtb += flux_adj(W[i], Wt[i], dt, Fb, Wb, Wtb);
        Q a, b, c, d;
        grp_adj(fp[i - 1], fm[i], s[i - 1], s[i], G, Wb, Wtb, a, b, c, d);
        fpb[i - 1] = Q{fpb[i - 1].r + a.r, fpb[i - 1].p + a.p, fpb[i - 1].un + a.un, fpb[i - 1].ut + a.ut};
        fmb[i] = Q{fmb[i].r + b.r, fmb[i].p + b.p, fmb[i].un + b.un, fmb[i].ut + b.ut};
        sb[i - 1] = Q{sb[i - 1].r + c.r, sb[i - 1].p + c.p, sb[i - 1].un + c.un, sb[i - 1].ut + c.ut};
        sb[i] = Q{sb[i].r + d.r, sb[i].p + d.p, sb[i].un + d.un, sb[i].ut + d.ut};
    }
    for (int k = 1; k <= n - 2; k++) {
        qb[k] = Q{qb[k].r + fmb[k].r + fpb[k].r, qb[k].p + fmb[k].p + fpb[k].p,
                  qb[k].un + fmb[k].un + fpb[k].un, qb[k].ut + fmb[k].ut + fpb[k].ut};
        const Q st{sb[k].r + L.hh * (fpb[k].r - fmb[k].r), sb[k].p + L.hh * (fpb[k].p - fmb[k].p),
                   sb[k].un + L.hh * (fpb[k].un - fmb[k].un), sb[k].ut + L.hh * (fpb[k].ut - fmb[k].ut)};
        Q pL, pC, pR;
        slope_adj(q[k - 1], q[k], q[k + 1], L, st, pL, pC, pR);
        auto add = [](Q &x, const Q &y) { x = Q{x.r + y.r, x.p + y.p, x.un + y.un, x.ut + y.ut}; };
        add(qb[k - 1], pL);
        add(qb[k], pC);
        add(qb[k + 1], pR);
    }
}

// X sweep of every row / Y sweep of every column (line order k = ny-1-j,
// normal velocity v), as the reference's rotated sweep orders it.
void sweep_x(const Mesh &m, const Field &U, Field &out, double dt) {
    out = U;
    std::vector<Q> line(m.nx), res;
    for (int j = 0; j < m.ny; j++) {
        for (int i = 0; i < m.nx; i++) line[i] = at(U, m, j, i);
        line_fwd(line, res, m.lx, m.gas, dt);
        for (int i = 0; i < m.nx; i++) at(out, m, j, i) = res[i];
    }
}
void sweep_y(const Mesh &m, const Field &U, Field &out, double dt) {
    out = U;
    std::vector<Q> line(m.ny), res;
    for (int i = 0; i < m.nx; i++) {
        for (int k = 0; k < m.ny; k++) line[k] = swapq(at(U, m, m.ny - 1 - k, i));
        line_fwd(line, res, m.ly, m.gas, dt);
        for (int k = 0; k < m.ny; k++) at(out, m, m.ny - 1 - k, i) = swapq(res[k]);
    }
}
void sweep_x_adj(const Mesh &m, const Field &U, const Field &ob, Field &ib, double dt, double &dtb) {
    ib.assign(U.size(), Q{0, 0, 0, 0});
    std::vector<Q> q(m.nx), qbo(m.nx), qb;
    for (int j = 0; j < m.ny; j++) {
        for (int i = 0; i < m.nx; i++) {
            q[i] = at(U, m, j, i);
            qbo[i] = at(ob, m, j, i);
        }
        line_adj(q, qbo, qb, m.lx, m.gas, dt, dtb);
        for (int i = 0; i < m.nx; i++) at(ib, m, j, i) = qb[i];
    }
}
void sweep_y_adj(const Mesh &m, const Field &U, const Field &ob, Field &ib, double dt, double &dtb) {
    ib.assign(U.size(), Q{0, 0, 0, 0});
    std::vector<Q> q(m.ny), qbo(m.ny), qb;
    for (int i = 0; i < m.nx; i++) {
        for (int k = 0; k < m.ny; k++) {
            q[k] = swapq(at(U, m, m.ny - 1 - k, i));
            qbo[k] = swapq(at(ob, m, m.ny - 1 - k, i));
        }
        line_adj(q, qbo, qb, m.ly, m.gas, dt, dtb);
        for (int k = 0; k < m.ny; k++) at(ib, m, m.ny - 1 - k, i) = swapq(qb[k]);
    }
}

// The preset's two phases as lists of (target, source, negate v) copies and
// (target, constant) writes; row phase first, then column phase.
struct Copy {
    int tj, ti, sj, si;
    bool negv;
};
struct Const {
    int tj, ti;
    Q v;
};

void fill_ops(const Mesh &m, double t, std::vector<Copy> ph[2], std::vector<Const> cs[2]) {
    const int Nx = m.nx, Ny = m.ny;
    const Q post{8, 116.5, 4.125 * std::sqrt(3.0), -4.125}, pre{1.4, 1, 0, 0};
    for (int p = 0; p < 2; p++) {
        ph[p].clear();
        cs[p].clear();
    }
    for (int j = 0; j < Ny; j++)
        for (int i = 0; i < Nx; i++) {
            if (m.strategy == 0 || m.strategy == 2 || m.strategy == 3) {
                if (j == 0) ph[0].push_back({j, i, 2, i, false});
                else if (j == 1) ph[0].push_back({j, i, 2, i, false});
                else if (j == Ny - 2) ph[0].push_back({j, i, Ny - 3, i, false});
                else if (j == Ny - 1) ph[0].push_back({j, i, m.strategy == 3 ? Ny - 4 : Ny - 3, i, false});
            } else {
                if (j == 0 || j == 1) {
                    const double temp = double(Ny) / std::sqrt(3.0), dist = double(Ny) * 20 * t / std::sqrt(3.0);
                    cs[0].push_back({j, i, i <= Nx / 24 + int(temp) + int(dist) ? post : pre});
                } else if (j == Ny - 1) {
                    ph[0].push_back({j, i, i <= Nx / 24 ? Ny - 3 : Ny - 4, i, i > Nx / 24});
                } else if (j == Ny - 2) {
                    ph[0].push_back({j, i, Ny - 3, i, i > Nx / 24});
                }
            }
            if (m.strategy == 0) {
                if (i == 0 || i == 1) ph[1].push_back({j, i, j, 2, false});
                else if (i == Nx - 2 || i == Nx - 1) ph[1].push_back({j, i, j, Nx - 3, false});
            } else if (m.strategy == 2) {
                if (i == 0 || i == 1)
                    cs[1].push_back({j, i, (j >= 2 * Ny / 5 + 1 && j <= 3 * Ny / 5) ? Q{5.0, 0.4127, 800.0, 0.0}
                                                                                  : Q{0.5, 0.4127, 0.0, 0.0}});
                else if (i == Nx - 2 || i == Nx - 1) ph[1].push_back({j, i, j, Nx - 3, false});
            } else if (m.strategy == 3) {
                if (i == 0 || i == 1) cs[1].push_back({j, i, j >= 9 * Ny / 10 ? Q{1.4, 1, 800, 0} : Q{0.14, 1, 0, 0}});
                else if (i == Nx - 2 || i == Nx - 1) ph[1].push_back({j, i, j, Nx - 3, false});
            } else {
                if (i == 0 || i == 1) cs[1].push_back({j, i, post});
                else if (i == Nx - 2 || i == Nx - 1) cs[1].push_back({j, i, pre});
            }
        }
}

void fill(const Mesh &m, Field &U, double t) {
    std::vector<Copy> ph[2];
    std::vector<Const> cs[2];
    fill_ops(m, t, ph, cs);
    for (int p = 0; p < 2; p++) {
        std::vector<Q> vals;
        for (auto &c : ph[p]) {
            Q v = at(U, m, c.sj, c.si);
            if (c.negv) v.ut = -v.ut;
            vals.push_back(v);
        }
        for (size_t k = 0; k < ph[p].size(); k++) at(U, m, ph[p][k].tj, ph[p][k].ti) = vals[k];
        for (auto &c : cs[p]) at(U, m, c.tj, c.ti) = c.v;
    }
}

// Reverse of the two phases: column phase first, then row phase.  A copy
// y <- x adds ybar to xbar and clears ybar; a constant write clears ybar.
void fill_adj(const Mesh &m, Field &lam) {
    std::vector<Copy> ph[2];
    std::vector<Const> cs[2];
    fill_ops(m, 0.0, ph, cs);
    for (int p = 1; p >= 0; p--) {
        for (auto &c : cs[p]) at(lam, m, c.tj, c.ti) = Q{0, 0, 0, 0};
        std::vector<Q> yb;
        for (auto &c : ph[p]) {
            yb.push_back(at(lam, m, c.tj, c.ti));
            at(lam, m, c.tj, c.ti) = Q{0, 0, 0, 0};
        }
        for (size_t k = 0; k < ph[p].size(); k++) {
            const Copy &c = ph[p][k];
            Q &x = at(lam, m, c.sj, c.si);
            x = Q{x.r + yb[k].r, x.p + yb[k].p, x.un + yb[k].un, x.ut + (c.negv ? -yb[k].ut : yb[k].ut)};
        }
    }
}

double smax_of(const Mesh &m, const Field &U, long long *argmax) {
    double best = -1.0;
    long long bi = 0;
    for (long long k = 0; k < (long long)U.size(); k++) {
        const double s = speed<false>(U[k].r, U[k].p, U[k].un, U[k].ut, m.gas);
        if (s > best) {
            best = s;
            bi = k;
        }
    }
    if (argmax) *argmax = bi;
    return best;
}

void step(const Mesh &m, const Field &U, Field &out, double t, double *dt_out) {
    const double smax = smax_of(m, U, nullptr);
    const double dt = m.cfl * m.hmin / smax;
    Field A1, A, B1, B;
    sweep_x(m, U, A1, dt);
    sweep_y(m, A1, A, dt);
    sweep_y(m, U, B1, dt);
    sweep_x(m, B1, B, dt);
    out.resize(U.size());
    for (size_t k = 0; k < U.size(); k++)
        out[k] = Q{(A[k].r + B[k].r) * 0.5, (A[k].p + B[k].p) * 0.5, (A[k].un + B[k].un) * 0.5,
                   (A[k].ut + B[k].ut) * 0.5};
    fill(m, out, t + dt);
    if (dt_out) *dt_out = dt;
}

// lam (dL/dU') -> dL/dU for the step taken from U.
void step_adj(const Mesh &m, const Field &U, const Field &lam_next, Field &lam) {
    long long am = 0;
    const double smax = smax_of(m, U, &am);
    const double dt = m.cfl * m.hmin / smax;
    Field A1, B1;
    sweep_x(m, U, A1, dt);
    sweep_y(m, U, B1, dt);
    Field lc = lam_next;
    fill_adj(m, lc);
    Field lh(lc.size());
    for (size_t k = 0; k < lc.size(); k++) lh[k] = Q{0.5 * lc[k].r, 0.5 * lc[k].p, 0.5 * lc[k].un, 0.5 * lc[k].ut};
    double dtb = 0.0;
    Field lA1, lB1, la, lb;
    sweep_y_adj(m, A1, lh, lA1, dt, dtb);
    sweep_x_adj(m, B1, lh, lB1, dt, dtb);
    sweep_x_adj(m, U, lA1, la, dt, dtb);
    sweep_y_adj(m, U, lB1, lb, dt, dtb);
    lam.resize(U.size());
    for (size_t k = 0; k < U.size(); k++)
        lam[k] = Q{la[k].r + lb[k].r, la[k].p + lb[k].p, la[k].un + lb[k].un, la[k].ut + lb[k].ut};
    const Q g = speed_grad(U[am].r, U[am].p, U[am].un, U[am].ut, m.gas);
    const double f = dtb * (-dt / smax);
    lam[am] = Q{lam[am].r + f * g.r, lam[am].p + f * g.p, lam[am].un + f * g.un, lam[am].ut + f * g.ut};
}

Mesh make_mesh(int nx, int ny, const orc_params *p) {
    Mesh m;
    m.nx = nx;
    m.ny = ny;
    m.lx = make_line(p->width / double(nx), p->alpha);
    m.ly = make_line(p->height / double(ny), p->alpha);
    m.gas = make_gas(p->gamma);
    m.cfl = p->cfl;
    m.hmin = std::min(p->width / double(nx), p->height / double(ny));
    m.strategy = p->strategy;
    return m;
}

// conserved planes (rho, rho u, rho v, E) <-> Field
void from_cons(const Mesh &m, const double *c, Field &U) {
    const size_t n = (size_t)m.nx * m.ny;
    U.resize(n);
    for (size_t k = 0; k < n; k++) {
        const double r = c[k], mu = c[n + k], mv = c[2 * n + k], E = c[3 * n + k];
        const double u = mu / r, v = mv / r;
        U[k] = Q{r, m.gas.g1 * (E - 0.5 * (mu * u + mv * v)), u, v};
    }
}
void to_cons(const Mesh &m, const Field &U, double *c) {
    const size_t n = (size_t)m.nx * m.ny;
    for (size_t k = 0; k < n; k++) {
        const Q &q = U[k];
        c[k] = q.r;
        c[n + k] = q.r * q.un;
        c[2 * n + k] = q.r * q.ut;
        c[3 * n + k] = q.p * m.gas.g1i + 0.5 * q.r * (q.un * q.un + q.ut * q.ut);
    }
}

}  // namespace

extern "C" {

// Forward of the differentiated solve: nsteps steps from U0 (conserved planes).
int orc_adj_forward(const double *U0cons, int nx, int ny, const orc_params *p, int nsteps, double *UNcons) {
    const Mesh m = make_mesh(nx, ny, p);
    Field U, V;
    from_cons(m, U0cons, U);
    double t = 0.0;
    for (int s = 0; s < nsteps; s++) {
        double dt = 0.0;
        step(m, U, V, t, &dt);
        t += dt;
        U.swap(V);
    }
    to_cons(m, U, UNcons);
    return 0;
}

// dL/dU0 (conserved) from dL/dU_N (conserved) by the reverse of every step.
int orc_adj_backward(const double *U0cons, int nx, int ny, const orc_params *p, int nsteps,
                     const double *gNcons, double *g0cons) {
    const Mesh m = make_mesh(nx, ny, p);
    std::vector<Field> states(nsteps + 1);
    from_cons(m, U0cons, states[0]);
    double t = 0.0;
    for (int s = 0; s < nsteps; s++) {
        double dt = 0.0;
        step(m, states[s], states[s + 1], t, &dt);
        t += dt;
    }
    const size_t n = (size_t)nx * ny;
    Field lam(n), lp;
    const Field &UN = states[nsteps];
    for (size_t k = 0; k < n; k++) {  // (d cons / d prim)^T
        const Q &q = UN[k];
        const double gr = gNcons[k], gmu = gNcons[n + k], gmv = gNcons[2 * n + k], gE = gNcons[3 * n + k];
        lam[k] = Q{gr + q.un * gmu + q.ut * gmv + 0.5 * (q.un * q.un + q.ut * q.ut) * gE, m.gas.g1i * gE,
                   q.r * gmu + q.r * q.un * gE, q.r * gmv + q.r * q.ut * gE};
    }
    for (int s = nsteps - 1; s >= 0; s--) {
        step_adj(m, states[s], lam, lp);
        lam.swap(lp);
    }
    for (size_t k = 0; k < n; k++) {  // (d prim / d cons)^T at U0
        const double r = U0cons[k], u = U0cons[n + k] / r, v = U0cons[2 * n + k] / r;
        const Q &l = lam[k];
        g0cons[k] = l.r - (l.un * u + l.ut * v) / r + l.p * m.gas.g1 * 0.5 * (u * u + v * v);
        g0cons[n + k] = l.un / r - l.p * m.gas.g1 * u;
        g0cons[2 * n + k] = l.ut / r - l.p * m.gas.g1 * v;
        g0cons[3 * n + k] = l.p * m.gas.g1;
    }
    return 0;
}

}  // extern "C"
