// Host build of the product's per-cell numerics (fvm2d_cell.h) and their
// hand-written adjoints (fvm2d_adjoint_cell.h) for CPU unit tests.
// Built by tests/test_cell_math.py with g++ -ffp-contract=off.
#include "../../2d_fvm_cuda_amd/csrc/fvm2d_adjoint_cell.h"

using namespace fvm2d;

static Q q4(const double *a) { return Q{a[0], a[1], a[2], a[3]}; }
static void put4(const Q &q, double *o) { o[0] = q.r; o[1] = q.p; o[2] = q.un; o[3] = q.ut; }
static void put5(const IfState &w, double *o) { o[0] = w.r; o[1] = w.p; o[2] = w.un; o[3] = w.ut; o[4] = w.e; }
static IfState s5(const double *a) { return IfState{a[0], a[1], a[2], a[3], a[4]}; }

extern "C" {
int cm_grp(int exact, const double *fl, const double *fr, const double *sl, const double *sr,
           double gamma, double *W, double *Wt) {
    Gas G = make_gas(gamma);
    IfState a, b;
    int br = exact ? grp<true>(q4(fl), q4(fr), q4(sl), q4(sr), G, a, b)
                   : grp<false>(q4(fl), q4(fr), q4(sl), q4(sr), G, a, b);
    put5(a, W);
    put5(b, Wt);
    return br;
}
void cm_flux(int exact, const double *W, const double *Wt, double k, double *F) {
    put4(exact ? flux<true>(s5(W), s5(Wt), k) : flux<false>(s5(W), s5(Wt), k), F);
}
void cm_slope(int exact, const double *qL, const double *q, const double *qR, double h, double alpha,
              double *s) {
    Line L = make_line(h, alpha);
    put4(exact ? slope<true>(q4(qL), q4(q), q4(qR), L) : slope<false>(q4(qL), q4(q), q4(qR), L), s);
}
void cm_update(int exact, const double *q, const double *FL, const double *FR, double lam, double gamma,
               double *out) {
    Gas G = make_gas(gamma);
    put4(exact ? update<true>(q4(q), q4(FL), q4(FR), lam, G) : update<false>(q4(q), q4(FL), q4(FR), lam, G),
         out);
}
double cm_speed(int exact, const double *q, double gamma) {
    Gas G = make_gas(gamma);
    return exact ? speed<true>(q[0], q[1], q[2], q[3], G) : speed<false>(q[0], q[1], q[2], q[3], G);
}
void cm_grp_adj(const double *fl, const double *fr, const double *sl, const double *sr, double gamma,
                const double *Wb, const double *Wtb, double *flb, double *frb, double *slb, double *srb) {
    Gas G = make_gas(gamma);
    Q a, b, c, d;
    grp_adj(q4(fl), q4(fr), q4(sl), q4(sr), G, s5(Wb), s5(Wtb), a, b, c, d);
    put4(a, flb);
    put4(b, frb);
    put4(c, slb);
    put4(d, srb);
}
double cm_flux_adj(const double *W, const double *Wt, double k, const double *Fb, double *Wb, double *Wtb) {
    IfState a, b;
    double kb = flux_adj(s5(W), s5(Wt), k, q4(Fb), a, b);
    put5(a, Wb);
    put5(b, Wtb);
    return kb;
}
void cm_slope_adj(const double *qL, const double *q, const double *qR, double h, double alpha,
                  const double *sb, double *qLb, double *qb, double *qRb) {
    Line L = make_line(h, alpha);
    Q a, b, c;
    slope_adj(q4(qL), q4(q), q4(qR), L, q4(sb), a, b, c);
    put4(a, qLb);
    put4(b, qb);
    put4(c, qRb);
}
void cm_prim_adj(const double *qo, const double *qob, double gamma, double *Ub) {
    put4(prim_adj(q4(qo), q4(qob), make_gas(gamma)), Ub);
}
void cm_cons_adj(const double *q, const double *Ub, double gamma, double *qb) {
    put4(cons_adj(q4(q), q4(Ub), make_gas(gamma)), qb);
}
void cm_speed_grad(const double *q, double gamma, double *g) {
    put4(speed_grad(q[0], q[1], q[2], q[3], make_gas(gamma)), g);
}
}
