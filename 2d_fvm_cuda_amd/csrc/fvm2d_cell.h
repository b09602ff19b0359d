// Per-cell and per-interface numerics of the relaxation-GRP MUSCL scheme.
//
// Everything works on a "line state" Q = (rho, p, un, ut): the density, the
// pressure, the velocity normal to the interfaces of the sweep and the
// tangential velocity.  An X-sweep uses (un, ut) = (u, v); a Y-sweep uses
// (un, ut) = (v, u), which is exactly the u<->v swap the reference applies
// with rotate_and_flip_vertically (src/basicFuns.cuh:249-262) before running
// its X kernels on the rotated array.
//
// Two instantiations:
//   Exact = true  : the reference's expressions with its evaluation order
//                   (src/basicFuns.cuh:14-212); compiled without FMA
//                   contraction it reproduces the host FP64 restatement bit
//                   for bit.
//   Exact = false : the same mathematics with reciprocals shared between
//                   terms (3 reciprocals + 2 reciprocal square roots per
//                   interface instead of ~62 divisions + 4 square roots),
//                   FMAs and a sign-test-free limiter.  Results differ from
//                   Exact only by rounding.
#pragma once

#include <cmath>

#if defined(__HIPCC__)
#include <hip/hip_runtime.h>
#define FVM_HD __host__ __device__ __forceinline__
#else
#define FVM_HD inline
#endif

namespace fvm2d {

struct Q {
    double r, p, un, ut;
};

// Interface record of the GRP: state (rho, p, un, ut, e) and its time
// derivative; e is the specific internal energy.
struct IfState {
    double r, p, un, ut, e;
};

struct Gas {
    double gamma;  // adiabatic exponent
    double g1;     // gamma - 1
    double g1i;    // 1 / (gamma - 1)
    double kappa;  // (gamma + 1) / 2
};

FVM_HD Gas make_gas(double gamma) {
    Gas g;
    g.gamma = gamma;
    g.g1 = gamma - 1.0;
    g.g1i = 1.0 / (gamma - 1.0);
    g.kappa = (gamma + 1.0) / 2.0;
    return g;
}

// Geometry of one sweep direction: h is the cell size along the sweep.
struct Line {
    double h;
    double ih;     // 1.0 / h
    double ih_a;   // (1.0 / h) * alpha
    double ih_c;   // 0.5 * (1.0 / h)
    double hh;     // 0.5 * h
    double alpha;
    double quarter;  // 0.25 as a register operand: it is no inline constant of
                     // the FP64 FMA, and as a literal it forces a move per use
};

FVM_HD Line make_line(double h, double alpha) {
    Line L;
    L.h = h;
    L.ih = 1.0 / h;
    L.ih_a = L.ih * alpha;
    L.ih_c = 0.5 * L.ih;
    L.hh = 0.5 * h;
    L.alpha = alpha;
    L.quarter = 0.25;
    return L;
}

// Field-wise select (a selected struct reference would live in private memory).
FVM_HD Q pick(bool c, const Q &a, const Q &b) {
    return Q{c ? a.r : b.r, c ? a.p : b.p, c ? a.un : b.un, c ? a.ut : b.ut};
}

FVM_HD double tmin(double a, double b) { return b < a ? b : a; }
FVM_HD double tmax(double a, double b) { return a < b ? b : a; }

// ---------------------------------------------------------------- limiter
// src/basicFuns.cuh:49-68: generalised minmod of the central difference
// (weight 1/2) and the one-sided differences (weight alpha).
template <bool Exact>
FVM_HD double limit1(double qL, double q, double qR, const Line &L) {
    const double c = (qR - qL) * 0.5 * L.ih;
    const double l = (q - qL) * L.ih * L.alpha;
    const double r = (qR - q) * L.ih * L.alpha;
    double s13 = c * r, s23 = l * r;
    if (s13 > 0.0 && s23 > 0.0) {
        double sgn = 1.0;
        if (c <= 0.0) sgn = -1.0;
        return sgn * tmin(fabs(c), tmin(fabs(l), fabs(r)));
    }
    return 0.0;
}

// Fast form of the same limiter.  With dC = qR - qL and sa = alpha carrying
// the sign of dC, min(sa dL, sa dR, |dC|/2) is positive exactly when dL and
// dR share the sign of dC (both one-sided differences agree), so clamping
// it at zero replaces the two sign tests; the scale 1/h is applied last.
template <>
FVM_HD double limit1<false>(double qL, double q, double qR, const Line &L) {
    const double dL = q - qL, dR = qR - q, dC = qR - qL;
    const double sa = copysign(L.alpha, dC);
    const double m = fmax(0.0, fmin(fmin(sa * dL, sa * dR), 0.5 * fabs(dC)));
    return m * copysign(L.ih, dC);
}

template <bool Exact>
FVM_HD Q slope(const Q &qL, const Q &q, const Q &qR, const Line &L) {
    return Q{limit1<Exact>(qL.r, q.r, qR.r, L), limit1<Exact>(qL.p, q.p, qR.p, L),
             limit1<Exact>(qL.un, q.un, qR.un, L), limit1<Exact>(qL.ut, q.ut, qR.ut, L)};
}

// src/basicFuns.cuh:69-70: face states q -/+ sigma*h/2.
template <bool Exact>
FVM_HD Q face_minus(const Q &q, const Q &s, const Line &L) {
    if (Exact)
        return Q{q.r - s.r * L.h * 0.5, q.p - s.p * L.h * 0.5, q.un - s.un * L.h * 0.5,
                 q.ut - s.ut * L.h * 0.5};
    return Q{fma(-s.r, L.hh, q.r), fma(-s.p, L.hh, q.p), fma(-s.un, L.hh, q.un),
             fma(-s.ut, L.hh, q.ut)};
}

template <bool Exact>
FVM_HD Q face_plus(const Q &q, const Q &s, const Line &L) {
    if (Exact)
        return Q{q.r + s.r * L.h * 0.5, q.p + s.p * L.h * 0.5, q.un + s.un * L.h * 0.5,
                 q.ut + s.ut * L.h * 0.5};
    return Q{fma(s.r, L.hh, q.r), fma(s.p, L.hh, q.p), fma(s.un, L.hh, q.un),
             fma(s.ut, L.hh, q.ut)};
}

// Fast-path slopes as limited differences across the cell, d = 2 h sigma
// = max(0, min(2 alpha dL, 2 alpha dR, |dC|)) with the sign of dC: the
// limiter without its 1/2 and 1/h multiplications, face states q -/+ d/4.
// The GRP's time derivatives are linear in the slopes, so with d in place
// of sigma they come out multiplied by 2h, and the flux takes k/(2h) in
// place of k.
FVM_HD double limit1_d(double qL, double q, double qR, double alpha2) {  // alpha2 = 2 alpha
    const double dL = q - qL, dR = qR - q, dC = qR - qL;
    const double sa = copysign(alpha2, dC);
    const double m = fmax(0.0, fmin(fmin(sa * dL, sa * dR), fabs(dC)));
    return copysign(m, dC);
}

FVM_HD Q slope_d(const Q &qL, const Q &q, const Q &qR, double alpha2) {
    return Q{limit1_d(qL.r, q.r, qR.r, alpha2), limit1_d(qL.p, q.p, qR.p, alpha2),
             limit1_d(qL.un, q.un, qR.un, alpha2), limit1_d(qL.ut, q.ut, qR.ut, alpha2)};
}

FVM_HD Q face_minus_d(const Q &q, const Q &d, double quarter) {
    return Q{fma(-quarter, d.r, q.r), fma(-quarter, d.p, q.p), fma(-quarter, d.un, q.un),
             fma(-quarter, d.ut, q.ut)};
}

FVM_HD Q face_plus_d(const Q &q, const Q &d, double quarter) {
    return Q{fma(quarter, d.r, q.r), fma(quarter, d.p, q.p), fma(quarter, d.un, q.un),
             fma(quarter, d.ut, q.ut)};
}

// Slopes, face states and the flux's time step of the step kernel's
// sweeps: the reference's forms in EXACT mode, cell units in FAST mode.
template <bool Exact>
FVM_HD Q sweep_slope(const Q &qL, const Q &q, const Q &qR, const Line &L) {
    if constexpr (Exact)
        return slope<true>(qL, q, qR, L);
    else
        return slope_d(qL, q, qR, 2.0 * L.alpha);
}

template <bool Exact>
FVM_HD Q sweep_face_minus(const Q &q, const Q &s, const Line &L) {
    if constexpr (Exact)
        return face_minus<true>(q, s, L);
    else
        return face_minus_d(q, s, L.quarter);
}

template <bool Exact>
FVM_HD Q sweep_face_plus(const Q &q, const Q &s, const Line &L) {
    if constexpr (Exact)
        return face_plus<true>(q, s, L);
    else
        return face_plus_d(q, s, L.quarter);
}

template <bool Exact>
FVM_HD double sweep_k(double dt, const Line &L) {
    return Exact ? dt : dt * L.ih_c;  // dt / (2h)
}

// ---------------------------------------------------------------- GRP
// Relaxation GRP at one interface, src/basicFuns.cuh:86-189.
//   fl, fr : left/right face states; sl, sr : slopes of the left/right cells.
// Returns the branch taken: 0 left-supersonic, 1 right-supersonic,
// 2 star region with u* <= 0, 3 star region with u* > 0.
template <bool Exact>
FVM_HD int grp(const Q &fl, const Q &fr, const Q &sl, const Q &sr, const Gas &G, IfState &W,
               IfState &Wt);

template <>
FVM_HD int grp<true>(const Q &fl, const Q &fr, const Q &sl, const Q &sr, const Gas &G,
                     IfState &W, IfState &Wt) {
    const double gamma = G.gamma;
    double rho_l = fl.r, rho_r = fr.r, pi_l = fl.p, pi_r = fr.p;
    double u_l = fl.un, u_r = fr.un, v_l = fl.ut, v_r = fr.ut;
    double e_l = pi_l / rho_l / (gamma - 1.0);
    double e_r = pi_r / rho_r / (gamma - 1.0);
    double c_l = 0.0, c_r = 0.0;
    double rho_lx = sl.r, rho_rx = sr.r, pi_lx = sl.p, pi_rx = sr.p;
    double u_lx = sl.un, u_rx = sr.un, v_lx = sl.ut, v_rx = sr.ut;
    double e_lx = pi_lx / rho_l / (gamma - 1.0) - pi_l * rho_lx / rho_l / rho_l / (gamma - 1.0);
    double e_rx = pi_rx / rho_r / (gamma - 1.0) - pi_r * rho_rx / rho_r / rho_r / (gamma - 1.0);
    double c_lx = 0.5 * sqrt(gamma / pi_l / rho_l) * (pi_l * rho_lx + rho_l * pi_lx);
    double c_rx = 0.5 * sqrt(gamma / pi_r / rho_r) * (pi_r * rho_rx + rho_r * pi_rx);
    double rho = 0.0, pi = 0.0, u = 0.0, v = 0.0, e = 0.0;
    double rho_t = 0.0, pi_t = 0.0, u_t = 0.0, v_t = 0.0, e_t = 0.0;
    double a_l = sqrt(gamma * pi_l / rho_l), a_r = sqrt(gamma * pi_r / rho_r);
    double kappa = (gamma + 1.0) / 2.0;
    if (pi_r >= pi_l) {
        double c_l_ = rho_l * (a_l + kappa * tmax(0.0, (pi_r - pi_l) / rho_r / a_r + u_l - u_r));
        c_l = c_l_;
        c_r = rho_r * (a_r + kappa * tmax(0.0, (pi_l - pi_r) / c_l_ + u_l - u_r));
    } else {
        double c_r_ = rho_r * (a_r + kappa * tmax(0.0, (pi_l - pi_r) / rho_l / a_l + u_l - u_r));
        double c_l_ = rho_l * (a_l + kappa * tmax(0.0, (pi_r - pi_l) / c_r_ + u_l - u_r));
        c_l = c_l_;
        c_r = c_r_;
    }
    int br;
    if (u_l - c_l / rho_l >= 0.0) {
        br = 0;
        rho = rho_l, pi = pi_l, u = u_l, v = v_l, e = e_l;
        rho_t = -rho * u_lx - u * rho_lx;
        u_t = -u * u_lx - pi_lx / rho;
        e_t = -u * e_lx - pi / rho * u_lx;
        v_t = -u * v_lx;
        pi_t = -u * pi_lx - c_l * c_l / rho_l * u_lx;
    } else if (u_r + c_r / rho_r <= 0.0) {
        br = 1;
        rho = rho_r, pi = pi_r, u = u_r, v = v_r, e = e_r;
        rho_t = -rho * u_rx - u * rho_rx;
        u_t = -u * u_rx - pi_rx / rho;
        e_t = -u * e_rx - pi / rho * u_rx;
        v_t = -u * v_rx;
        pi_t = -u * pi_rx - c_r * c_r / rho * u_rx;
    } else {
        double u_m = (u_l * c_l + u_r * c_r + pi_l - pi_r) / (c_l + c_r);
        double pi_m = ((u_l - u_r) * c_l * c_r + c_l * pi_r + c_r * pi_l) / (c_l + c_r);
        // u is still 0.0 in these four (src/basicFuns.cuh:110,151-154)
        double k_l = -pi_lx - c_l * u_lx + 0.5 * (u - u_l) * c_lx;
        double k_r = -pi_rx + c_r * u_rx - 0.5 * (u - u_r) * c_rx;
        double d_l = -2.0 * c_l * c_l / rho_l * u_lx - 2.0 * c_l / rho_l * pi_lx +
                     c_l / rho_l * (u - u_l) * c_lx;
        double d_r = 2.0 * c_r * c_r / rho_r * u_rx - 2.0 * c_r / rho_r * pi_rx +
                     c_r / rho_r * (u_r - u) * c_rx;
        double Dpi_Dt = (c_r * d_l - c_l * d_r) / 2.0 / (c_l + c_r);
        pi = pi_m;
        u = u_m;
        if (u <= 0.0) {
            br = 2;
            double rho_mr = c_r / (u_r - u_m + c_r / rho_r);
            double e_mr = e_r + (pi_m - pi_r) * (pi_m + pi_r) / 2.0 / c_r / c_r;
            rho = rho_mr;
            e = e_mr;
            v = v_r;
            rho_t = -u * rho * rho * rho / (rho_r * c_r * c_r) *
                        (c_r * c_r / rho_r / rho_r * rho_rx - c_r * u_rx + 1.5 * (u - u_r) * c_rx) +
                    rho * rho * rho / (c_r * c_r * c_r) * (u + c_r / rho) * Dpi_Dt;
            pi_t = 1.0 / (c_l + c_r) *
                   (c_l * c_r / rho_l * (1 + u * rho / c_r) * k_l -
                    c_l * c_r / rho_r * (1 - u * rho / c_l) * k_r);
            u_t = 1.0 / (c_l + c_r) *
                  (c_l / rho_l * (1 + u * rho / c_r) * k_l +
                   c_r / rho_r * (1 - u * rho * c_l / c_r / c_r) * k_r);
            e_t = -u * rho / rho_r / c_r / c_r / c_r *
                      (-pi_r * c_r * pi_rx + c_r * c_r * c_r * e_rx +
                       2.0 * (e_r - e) * c_r * c_r * c_rx) +
                  pi / c_r / c_r * pi_t;
            double v_x = rho / rho_r * v_rx;
            v_t = -u * v_x;
        } else {
            br = 3;
            double rho_ml = c_l / (u_m - u_l + c_l / rho_l);
            double e_ml = e_l + (pi_m - pi_l) * (pi_m + pi_l) / 2.0 / c_l / c_l;
            rho = rho_ml;
            e = e_ml;
            v = v_l;
            rho_t = -u * rho * rho * rho / (rho_l * c_l * c_l) *
                        (c_l * c_l / rho_l / rho_l * rho_lx + c_l * u_lx + 1.5 * (u - u_l) * c_lx) -
                    rho * rho * rho / (c_l * c_l * c_l) * (u - c_l / rho) * Dpi_Dt;
            pi_t = 1.0 / (c_l + c_r) *
                   (c_l * c_r / rho_l * (1 + u * rho / c_r) * k_l -
                    c_l * c_r / rho_r * (1 - u * rho / c_l) * k_r);
            u_t = 1.0 / (c_l + c_r) *
                  (c_l / rho_l * (1 + u * rho * c_r / c_l / c_l) * k_l +
                   c_r / rho_r * (1 - u * rho / c_l) * k_r);
            e_t = -u * rho / rho_l / c_l / c_l / c_l *
                      (-pi_l * c_l * pi_lx + c_l * c_l * c_l * e_lx +
                       2.0 * (e_l - e) * c_l * c_l * c_lx) +
                  pi / c_l / c_l * pi_t;
            double v_x = rho / rho_l * v_lx;
            v_t = -u * v_x;
        }
    }
    W = IfState{rho, pi, u, v, e};
    Wt = IfState{rho_t, pi_t, u_t, v_t, e_t};
    return br;
}

// Reciprocal of the fast path.  On the device: v_rcp_f64 (24 correct bits
// on gfx950, measured: max relative error 4.5e-8) refined by one Newton step
// r <- r + r (1 - x r), which squares the error (measured: 2.2e-15 over
// 2^20 arguments in [2^-8, 2^8]); on the host: IEEE division.
FVM_HD double rcp(double x) {
#if defined(__HIP_DEVICE_COMPILE__)
    const double r = __builtin_amdgcn_rcp(x);
    return fma(r, fma(-x, r, 1.0), r);
#else
    return 1.0 / x;
#endif
}

// Reciprocal square root of the fast path (positive normal x).  On the
// device: v_rsq_f64 (24 correct bits, measured max relative error 5.2e-8)
// refined by one Newton step y <- y + y (1/2 - (x/2) y^2) (measured: 4.1e-15);
// on the host: IEEE operations.
FVM_HD double rsq(double x) {
#if defined(__HIP_DEVICE_COMPILE__)
    const double y = __builtin_amdgcn_rsq(x);
    return fma(y, fma(-(0.5 * x) * y, y, 0.5), y);
#else
    return 1.0 / sqrt(x);
#endif
}

template <>
FVM_HD int grp<false>(const Q &fl, const Q &fr, const Q &sl, const Q &sr, const Gas &G,
                      IfState &W, IfState &Wt) {
    const double gamma = G.gamma, g1i = G.g1i, kappa = G.kappa;
    const double rl = fl.r, rr = fr.r, pl = fl.p, pr = fr.p;
    const double ul = fl.un, ur = fr.un;
    // Lagrangian sound speeds S = rho a = sqrt(gamma p rho) from one
    // reciprocal square root each; 1/rho = gamma p / S^2
    const double gpl = gamma * pl, gpr = gamma * pr;
    const double xl = gpl * rl, xr = gpr * rr;
    const double iSl = rsq(xl), iSr = rsq(xr);
    const double Sl = xl * iSl, Sr = xr * iSr;
    const double irl = gpl * iSl * iSl, irr = gpr * iSr * iSr;
    // c_x = 0.5 sqrt(gamma/(p rho)) (p rho_x + rho p_x) = (gamma/2) (...)/S;
    // clx and crx below leave out the factor gamma/2, which their uses take
    const double clx = iSl * fma(pl, sl.r, rl * sl.p);
    const double crx = iSr * fma(pr, sr.r, rr * sr.p);
    const double du = ul - ur;
    // Relaxation speeds (src/basicFuns.cuh:94-104).  The side F with the
    // lower pressure is relaxed first from the other side's S, then the
    // other side O from c_F; |p_r - p_l| is the pressure rise seen by F.
    const bool fL = pr >= pl;
    const double dp = fabs(pr - pl);
    const double SF = fL ? Sl : Sr, rF = fL ? rl : rr, iSO = fL ? iSr : iSl;
    const double SO = fL ? Sr : Sl, rO = fL ? rr : rl;
    const double cF = fma(rF * kappa, fmax(0.0, fma(dp, iSO, du)), SF);
    const double icF = rcp(cF);
    const double cO = fma(rO * kappa, fmax(0.0, fma(-dp, icF, du)), SO);
    // 1/c_O and 1/(c_l + c_r) from one reciprocal
    const double csum = cF + cO;
    const double t = rcp(cO * csum);
    const double icO = csum * t, ics = cO * t;
    const double cl = fL ? cF : cO, cr = fL ? cO : cF;
    const double icl = fL ? icF : icO, icr = fL ? icO : icF;
    const bool supL = fma(-cl, irl, ul) >= 0.0;  // u_l - c_l / rho_l >= 0
    const bool supR = !supL && fma(cr, irr, ur) <= 0.0;
    if (supL || supR) {
        // upwind state with the primitive Euler time derivatives; the two
        // supersonic branches differ only in which side K is upwind.  Kept
        // as a branch: most waves have no supersonic interface and skip it
        // (measured: 13% less time per step than evaluating both and selecting).
        const double rK = supL ? rl : rr, pK = supL ? pl : pr, u = supL ? ul : ur;
        const double irK = supL ? irl : irr, cK = supL ? cl : cr;
        const Q sK = pick(supL, sl, sr);
        // e = p / rho / (gamma - 1), e_x = (p_x - p rho_x / rho) / rho / (gamma - 1)
        const double prK = pK * irK;
        const double eK = prK * g1i, eKx = fma(-prK, sK.r, sK.p) * (irK * g1i);
        W = IfState{rK, pK, u, supL ? fl.ut : fr.ut, eK};
        Wt.r = -fma(rK, sK.un, u * sK.r);
        Wt.un = -fma(u, sK.un, sK.p * irK);
        Wt.e = -fma(u, eKx, pK * irK * sK.un);
        Wt.ut = -u * sK.ut;
        Wt.p = -fma(u, sK.p, cK * cK * irK * sK.un);
        return supL ? 0 : 1;
    }
    const double um = (fma(ul, cl, fma(ur, cr, pl - pr))) * ics;
    const double clcr = cl * cr;
    const double pm = fma(clcr, du, fma(cl, pr, cr * pl)) * ics;
    // The two star sub-cases (u* <= 0: right state K = r, s = -1; u* > 0:
    // left state K = l, s = +1) are mirror images of each other except for
    // the 1.5 (u* - u_K) c_Kx term, which keeps its sign in both.
    const bool right = um <= 0.0;
    const double s = right ? -1.0 : 1.0;
    const double cK = right ? cr : cl, icK = right ? icr : icl;
    const double irK = right ? irr : irl;
    const double uK = right ? ur : ul, pK = right ? pr : pl;
    const double cKx = right ? crx : clx;
    const Q sK = pick(right, sr, sl);
    // e and e_x of side K only (the star state needs no other)
    const double prK = pK * irK;
    const double eK = prK * g1i, eKx = fma(-prK, sK.r, sK.p) * (irK * g1i);
    // k_l, k_r use u = 0 (see the exact form).  The reference's d_K equal
    // 2 (c_K / rho_K) k_K, so with P = k_l / rho_l, R = k_r / rho_r and
    // M = (c_l P + c_r R) / (c_l + c_r) its three time derivatives of both
    // star sub-cases (src/basicFuns.cuh:151-185) reduce to
    //   Dpi/Dt = c_l c_r (P - R) / (c_l + c_r),
    //   p_t = Dpi/Dt + u* rho* M,   u_t = M + u* rho* Dpi/Dt / c_K^2.
    const double g4 = 0.25 * gamma;  // (1/2) (gamma/2)
    const double kl = fma(-g4 * ul, clx, fma(-cl, sl.un, -sl.p));
    const double kr = fma(g4 * ur, crx, fma(cr, sr.un, -sr.p));
    const double P = kl * irl, R = kr * irr;
    const double Dp = clcr * ics * (P - R);
    const double M = fma(cl, P, cr * R) * ics;
    const double cirK = cK * irK;
    const double den = fma(s, um - uK, cirK);  // c_K / rho*
    const double iden = rcp(den);
    const double rho = cK * iden;
    const double icK2 = icK * icK;
    // e* = e_K + (p*^2 - p_K^2) / (2 c_K^2)
    const double D2 = (pm - pK) * (pm + pK) * icK2;
    const double e = fma(0.5, D2, eK);
    // c_K^2/rho_K^2 rho_Kx + s c_K u_Kx + 1.5 (u* - u_K) c_Kx
    const double t1 = fma(cirK * cirK, sK.r, fma(s * cK, sK.un, (0.75 * gamma) * (um - uK) * cKx));
    // rho*^3 / (rho_K c_K^2) = (c_K / rho_K) / den^3 and rho*^3 / c_K^3 =
    // 1 / den^3; the reference's (c_K / rho* - s u*), with c_K / rho* = den
    // = s (u* - u_K) + c_K / rho_K, reduces to c_K / rho_K - s u_K
    const double iden3 = iden * iden * iden;
    const double umc = um * cirK;
    const double rho_t = iden3 * fma(-umc, t1, fma(-s, uK, cirK) * Dp);
    const double urho = um * rho;
    const double qK = umc * iden;  // u* rho* / rho_K
    const double p_t = fma(urho, M, Dp);
    const double u_t = fma(urho * icK2, Dp, M);
    // 2 (e_K - e*) / c_K = -D2 / c_K
    const double e_t = fma(-qK, fma(-(0.5 * gamma) * D2 * icK, cKx, fma(-pK * icK2, sK.p, eKx)),
                           pm * icK2 * p_t);
    W = IfState{rho, pm, um, right ? fr.ut : fl.ut, e};
    Wt = IfState{rho_t, p_t, u_t, -qK * sK.ut, e_t};
    return right ? 2 : 3;
}

// ---------------------------------------------------------------- flux
// F = f(W) + (k/2) d f(W)/dt in the order (mass, normal momentum, energy,
// tangential momentum), src/basicFuns.cuh:14-34; the energy flux uses the
// GRP's e, not p/((gamma-1) rho).
template <bool Exact>
FVM_HD Q flux(const IfState &U, const IfState &Ut, double k) {
    const double rho = U.r, pi = U.p, u = U.un, v = U.ut, e = U.e;
    const double rho_t = Ut.r, pi_t = Ut.p, u_t = Ut.un, v_t = Ut.ut, e_t = Ut.e;
    if (Exact) {
        return Q{rho * u + k / 2.0 * (rho * u_t + u * rho_t),
                 rho * u * u + pi + k / 2.0 * (rho_t * u * u + 2.0 * rho * u * u_t + pi_t),
                 rho * u * (e + u * u / 2.0 + v * v / 2.0) + pi * u +
                     k / 2.0 *
                         (rho_t * u * e + rho * u_t * e + rho * u * e_t + rho_t * u * u * u / 2.0 +
                          1.5 * rho * u * u * u_t + rho_t * u * v * v / 2.0 + rho * u * v * v_t +
                          rho * u_t * v * v / 2.0 + pi_t * u + pi * u_t),
                 rho * u * v + k / 2.0 * (rho * u_t * v + u * rho_t * v + u * rho * v_t)};
    }
    const double kh = 0.5 * k;
    const double ru = rho * u;
    const double K = fma(0.5, fma(u, u, v * v), e);  // e + (u^2 + v^2)/2
    const double H = fma(rho, K, pi);                  // rho K + p
    const double mt = fma(rho, u_t, u * rho_t);        // (rho u)_t
    const double f0 = fma(kh, mt, ru);
    // (rho u^2 + p)_t = u (rho u)_t + rho u u_t + p_t
    const double f1 = fma(kh, fma(u, mt, fma(ru, u_t, pi_t)), fma(ru, u, pi));
    // (u H)_t = u_t H + u ((rho K)_t + p_t), (rho K)_t = rho_t K + rho (e_t + u u_t + v v_t)
    const double g2 = fma(u_t, H, u * (fma(rho_t, K, rho * fma(u, u_t, fma(v, v_t, e_t))) + pi_t));
    const double f2 = fma(kh, g2, u * H);
    const double f3 = fma(kh, fma(v, mt, ru * v_t), ru * v);
    return Q{f0, f1, f2, f3};
}

// ---------------------------------------------------------------- update
// src/basicFuns.cuh:229-246 with src/basicFuns.cuh:195-212: convert to
// conserved (rho, rho un, E, rho ut), subtract lambda (F_right - F_left),
// convert back to primitive.
template <bool Exact>
FVM_HD Q update(const Q &q, const Q &FL, const Q &FR, double lambda, const Gas &G) {
    if (Exact) {
        const double gamma = G.gamma;
        double c0 = q.r, c1 = q.r * q.un,
               c2 = q.p / (gamma - 1.0) + q.r * (q.un * q.un + q.ut * q.ut) / 2.0, c3 = q.r * q.ut;
        c0 = c0 - (FR.r - FL.r) * lambda;
        c1 = c1 - (FR.p - FL.p) * lambda;
        c2 = c2 - (FR.un - FL.un) * lambda;
        c3 = c3 - (FR.ut - FL.ut) * lambda;
        return Q{c0, (gamma - 1.0) * (c2 - c1 * c1 / c0 / 2.0 - c3 * c3 / c0 / 2.0), c1 / c0, c3 / c0};
    }
    const double mu = q.r * q.un, mv = q.r * q.ut;  // momenta
    const double c0 = fma(-lambda, FR.r - FL.r, q.r);
    const double c1 = fma(-lambda, FR.p - FL.p, mu);
    const double E = fma(0.5, fma(mu, q.un, mv * q.ut), q.p * G.g1i);  // p/(g-1) + rho |u|^2/2
    const double c2 = fma(-lambda, FR.un - FL.un, E);
    const double c3 = fma(-lambda, FR.ut - FL.ut, mv);
    const double ir = rcp(c0);
    const double un = c1 * ir, ut = c3 * ir;
    const double p = G.g1 * fma(-0.5, fma(c1, un, c3 * ut), c2);  // (g-1) (E - rho |u|^2/2)
    return Q{c0, p, un, ut};
}

// src/basicFuns.cuh:223-224: |(u, v)| + sqrt(gamma p / rho).
template <bool Exact>
FVM_HD double speed(double rho, double p, double u, double v, const Gas &G) {
    if (Exact) {
        double velo = sqrt(u * u + v * v), c = sqrt(G.gamma * p / rho);
        return fabs(velo) + c;
    }
    // a = sqrt(gamma p / rho) = gamma p / sqrt(gamma p rho); |(u, v)| = q / sqrt(q)
    // (below 1e-300, q contributes nothing at double precision next to a)
    const double gp = G.gamma * p, q = fma(u, u, v * v);
    const double vel = q > 1e-300 ? q * rsq(q) : 0.0;
    return vel + gp * rsq(gp * rho);
}

}  // namespace fvm2d
