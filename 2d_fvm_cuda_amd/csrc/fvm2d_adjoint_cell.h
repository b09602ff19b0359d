// Hand-written reverse mode of the per-cell and per-interface numerics of
// fvm2d_cell.h (the FAST formulation; EXACT evaluates the same functions in
// another rounding order).  Every routine recomputes its forward
// intermediates, then walks the forward statements backwards, accumulating
// the adjoint ("bar") of each intermediate.  Branches (limiter selection,
// relaxation-speed max terms, supersonic / star sub-cases) are
// differentiated along the branch the forward evaluation takes.
#pragma once

#include "fvm2d_cell.h"

// The reverse statements are sums of products: on the device they may
// contract to fused multiply-adds (the file is not part of the bitwise
// EXACT path).  The host build of the reference gradient keeps IEEE
// operations; the two agree to rounding.
#if defined(__HIP_DEVICE_COMPILE__)
#pragma clang fp contract(fast)
#endif

namespace fvm2d {

// ---------------------------------------------------------------- limiter
// Adjoint of limit1<false>: sigma = the one of (c, l, r) with the smallest
// magnitude when all three share a sign, else 0.  Ties resolve to c, then l.
FVM_HD void limit1_adj(double qL, double q, double qR, const Line &L, double sb, double &qLb,
                       double &qb, double &qRb) {
    const double c = (qR - qL) * L.ih_c;
    const double l = (q - qL) * L.ih_a;
    const double r = (qR - q) * L.ih_a;
    // branch-free (selected weights): assigning the three outputs in
    // separate branches makes the compiler keep them in private memory
    const bool act = c * r > 0.0 && l * r > 0.0;
    const double ac = fabs(c), al = fabs(l), ar = fabs(r);
    const bool selc = ac <= al && ac <= ar;
    const bool sell = !selc && al <= ar;
    const bool selr = !selc && !sell;
    const double wL = selc ? -L.ih_c : (sell ? -L.ih_a : 0.0);
    const double wC = sell ? L.ih_a : (selr ? -L.ih_a : 0.0);
    const double wR = selc ? L.ih_c : (selr ? L.ih_a : 0.0);
    qLb = act ? wL * sb : 0.0;
    qb = act ? wC * sb : 0.0;
    qRb = act ? wR * sb : 0.0;
}

FVM_HD void slope_adj(const Q &qL, const Q &q, const Q &qR, const Line &L, const Q &sb, Q &qLb,
                      Q &qb, Q &qRb) {
    limit1_adj(qL.r, q.r, qR.r, L, sb.r, qLb.r, qb.r, qRb.r);
    limit1_adj(qL.p, q.p, qR.p, L, sb.p, qLb.p, qb.p, qRb.p);
    limit1_adj(qL.un, q.un, qR.un, L, sb.un, qLb.un, qb.un, qRb.un);
    limit1_adj(qL.ut, q.ut, qR.ut, L, sb.ut, qLb.ut, qb.ut, qRb.ut);
}

// ---------------------------------------------------------------- flux
// Adjoint of flux<false>(W, Wt, k) for the output adjoint Fb.  Returns
// dF/dk . Fb (the sensitivity to the time step k).
FVM_HD double flux_adj(const IfState &U, const IfState &Ut, double k, const Q &Fb, IfState &Ub,
                       IfState &Utb) {
    const double rho = U.r, p = U.p, u = U.un, v = U.ut, e = U.e;
    const double rt = Ut.r, pt = Ut.p, ut = Ut.un, vt = Ut.ut, et = Ut.e;
    const double b0 = Fb.r, b1 = Fb.p, b2 = Fb.un, b3 = Fb.ut;
    const double kh = 0.5 * k;
    const double ru = rho * u;
    const double K = e + 0.5 * u * u + 0.5 * v * v;
    const double mt = rho * ut + u * rt;
    const double h1 = rt * u * u + 2.0 * ru * ut + pt;
    const double g2 = rt * u * K + rho * ut * (K + u * u) + ru * (v * vt + et) + pt * u + p * ut;
    const double h3 = v * mt + ru * vt;
    const double khb = mt * b0 + h1 * b1 + g2 * b2 + h3 * b3;
    // reverse of f3 = ru v + kh (v mt + ru vt)
    double rub = v * b3 + kh * vt * b3;
    double vb = ru * b3 + kh * mt * b3;
    double mtb = kh * v * b3;
    double vtb = kh * ru * b3;
    // f2 = ru K + p u + kh g2
    double Kb = ru * b2;
    double pb = u * b2;
    double ub = p * b2;
    rub += K * b2;
    const double g2b = kh * b2;
    double rtb = u * K * g2b;
    ub += rt * K * g2b;
    Kb += rt * u * g2b;
    double rhob = ut * (K + u * u) * g2b;
    double utb = rho * (K + u * u) * g2b;
    Kb += rho * ut * g2b;
    ub += 2.0 * rho * ut * u * g2b;
    rub += (v * vt + et) * g2b;
    vb += ru * vt * g2b;
    vtb += ru * v * g2b;
    double etb = ru * g2b;
    double ptb = u * g2b;
    ub += pt * g2b;
    pb += ut * g2b;
    utb += p * g2b;
    // f1 = ru u + p + kh (rt u^2 + 2 ru ut + pt)
    rub += u * b1 + 2.0 * kh * ut * b1;
    ub += ru * b1 + 2.0 * kh * rt * u * b1;
    pb += b1;
    rtb += kh * u * u * b1;
    utb += 2.0 * kh * ru * b1;
    ptb += kh * b1;
    // f0 = ru + kh mt
    rub += b0;
    mtb += kh * b0;
    // mt = rho ut + u rt
    rhob += ut * mtb;
    utb += rho * mtb;
    ub += rt * mtb;
    rtb += u * mtb;
    // K = e + u^2/2 + v^2/2
    double eb = Kb;
    ub += u * Kb;
    vb += v * Kb;
    // ru = rho u
    rhob += u * rub;
    ub += rho * rub;
    Ub = IfState{rhob, pb, ub, vb, eb};
    Utb = IfState{rtb, ptb, utb, vtb, etb};
    return 0.5 * khb;
}

// ---------------------------------------------------------------- GRP
// Adjoint of grp<false>.  Outputs the adjoints of the left/right face
// states and of the left/right cell slopes.
FVM_HD void grp_adj(const Q &fl, const Q &fr, const Q &sl, const Q &sr, const Gas &G,
                    const IfState &Wb, const IfState &Wtb, Q &flb, Q &frb, Q &slb, Q &srb) {
    const double gamma = G.gamma, g1i = G.g1i, kappa = G.kappa;
    const double rl = fl.r, rr = fr.r, pl = fl.p, pr = fr.p, ul = fl.un, ur = fr.un;
    const double rlx = sl.r, plx = sl.p, ulx = sl.un, vlx = sl.ut;
    const double rrx = sr.r, prx = sr.p, urx = sr.un, vrx = sr.ut;
    // ---- forward (same statements as grp<false>)
    // S = sqrt(gamma p rho), 1/S and 1/rho = gamma p / S^2 from one
    // reciprocal square root per side (the reverse below uses only the
    // mathematical relations between these quantities)
    const double gpl = gamma * pl, gpr = gamma * pr;
    const double Xl = gpl * rl, Xr = gpr * rr;
    const double iSl = rsq(Xl), iSr = rsq(Xr);
    const double Sl = Xl * iSl, Sr = Xr * iSr;
    const double irl = gpl * iSl * iSl, irr = gpr * iSr * iSr;
    const double el = pl * irl * g1i, er = pr * irr * g1i;
    const double zl = plx - pl * irl * rlx, zr = prx - pr * irr * rrx;
    const double elx = zl * irl * g1i, erx = zr * irr * g1i;
    const double sxl = pl * rlx + rl * plx, sxr = pr * rrx + rr * prx;
    const double clx = 0.5 * gamma * iSl * sxl, crx = 0.5 * gamma * iSr * sxr;
    const double du = ul - ur;
    const bool P = pr >= pl;
    double ml, mr, xl, xr, cl, cr, icl, icr;
    if (P) {
        xl = (pr - pl) * iSr + du;
        ml = tmax(0.0, xl);
        cl = Sl + rl * kappa * ml;
        icl = rcp(cl);
        xr = (pl - pr) * icl + du;
        mr = tmax(0.0, xr);
        cr = Sr + rr * kappa * mr;
        icr = rcp(cr);
    } else {
        xr = (pl - pr) * iSl + du;
        mr = tmax(0.0, xr);
        cr = Sr + rr * kappa * mr;
        icr = rcp(cr);
        xl = (pr - pl) * icr + du;
        ml = tmax(0.0, xl);
        cl = Sl + rl * kappa * ml;
        icl = rcp(cl);
    }
    // ---- adjoints of intermediates
    double b_rl = 0, b_pl = 0, b_ul = 0, b_vl = 0, b_rr = 0, b_pr = 0, b_ur = 0, b_vr = 0;
    double b_rlx = 0, b_plx = 0, b_ulx = 0, b_vlx = 0, b_rrx = 0, b_prx = 0, b_urx = 0, b_vrx = 0;
    double b_irl = 0, b_irr = 0, b_el = 0, b_er = 0, b_elx = 0, b_erx = 0, b_clx = 0, b_crx = 0;
    double b_cl = 0, b_cr = 0, b_icl = 0, b_icr = 0, b_du = 0;

    if (ul - cl * irl >= 0.0) {  // left supersonic
        b_rl += Wb.r;
        b_pl += Wb.p;
        b_ul += Wb.un;
        b_vl += Wb.ut;
        b_el += Wb.e;
        // rho_t = -(rl ulx + ul rlx)
        b_rl -= ulx * Wtb.r;
        b_ulx -= rl * Wtb.r;
        b_ul -= rlx * Wtb.r;
        b_rlx -= ul * Wtb.r;
        // u_t = -(ul ulx + plx irl)
        b_ul -= ulx * Wtb.un;
        b_ulx -= ul * Wtb.un;
        b_plx -= irl * Wtb.un;
        b_irl -= plx * Wtb.un;
        // e_t = -(ul elx + pl irl ulx)
        b_ul -= elx * Wtb.e;
        b_elx -= ul * Wtb.e;
        b_pl -= irl * ulx * Wtb.e;
        b_irl -= pl * ulx * Wtb.e;
        b_ulx -= pl * irl * Wtb.e;
        // v_t = -ul vlx
        b_ul -= vlx * Wtb.ut;
        b_vlx -= ul * Wtb.ut;
        // p_t = -(ul plx + cl^2 irl ulx)
        b_ul -= plx * Wtb.p;
        b_plx -= ul * Wtb.p;
        b_cl -= 2.0 * cl * irl * ulx * Wtb.p;
        b_irl -= cl * cl * ulx * Wtb.p;
        b_ulx -= cl * cl * irl * Wtb.p;
    } else if (ur + cr * irr <= 0.0) {  // right supersonic
        b_rr += Wb.r;
        b_pr += Wb.p;
        b_ur += Wb.un;
        b_vr += Wb.ut;
        b_er += Wb.e;
        b_rr -= urx * Wtb.r;
        b_urx -= rr * Wtb.r;
        b_ur -= rrx * Wtb.r;
        b_rrx -= ur * Wtb.r;
        b_ur -= urx * Wtb.un;
        b_urx -= ur * Wtb.un;
        b_prx -= irr * Wtb.un;
        b_irr -= prx * Wtb.un;
        b_ur -= erx * Wtb.e;
        b_erx -= ur * Wtb.e;
        b_pr -= irr * urx * Wtb.e;
        b_irr -= pr * urx * Wtb.e;
        b_urx -= pr * irr * Wtb.e;
        b_ur -= vrx * Wtb.ut;
        b_vrx -= ur * Wtb.ut;
        b_ur -= prx * Wtb.p;
        b_prx -= ur * Wtb.p;
        b_cr -= 2.0 * cr * irr * urx * Wtb.p;
        b_irr -= cr * cr * urx * Wtb.p;
        b_urx -= cr * cr * irr * Wtb.p;
    } else {  // star region
        // ---- forward of the common star part
        const double ics = rcp(cl + cr);
        const double Nu = ul * cl + ur * cr + pl - pr;
        const double um = Nu * ics;
        const double Np = du * cl * cr + cl * pr + cr * pl;
        const double pm = Np * ics;
        const double Ql = -2.0 * cl * ulx - ul * clx - 2.0 * plx;
        const double Qr = 2.0 * cr * urx + ur * crx - 2.0 * prx;
        const double kl = -plx - cl * ulx - 0.5 * ul * clx;
        const double kr = -prx + cr * urx + 0.5 * ur * crx;
        const double dl = cl * irl * Ql, dr = cr * irr * Qr;
        const double ND = cr * dl - cl * dr;
        const double Dp = ND * 0.5 * ics;
        double b_um = Wb.un, b_pm = Wb.p, b_ics = 0, b_Dp = 0, b_kl = 0, b_kr = 0;
        const double b_rho_t = Wtb.r, b_v_t = Wtb.ut, b_e_t = Wtb.e, b_u_t = Wtb.un;
        double b_p_t = Wtb.p;
        double b_rho = Wb.r, b_e = Wb.e;
        if (um <= 0.0) {
            // ---- forward of the right sub-case
            const double den = ur - um + cr * irr;
            const double iden = rcp(den);
            const double rho = cr * iden;
            const double irho = den * icr;
            const double icr2 = icr * icr;
            const double dpm = pm - pr, spm = pm + pr;
            const double e = er + dpm * spm * 0.5 * icr2;
            const double r3 = rho * rho * rho;
            const double t1 = cr * cr * irr * irr * rrx - cr * urx + 1.5 * (um - ur) * crx;
            const double X1 = -um * r3 * irr * icr2;
            const double Z = um + cr * irho;
            const double X2 = r3 * icr2 * icr * Z;
            const double urho = um * rho;
            const double fA = 1.0 + urho * icr;
            const double A = cl * irl * fA * kl;
            const double fB = 1.0 - urho * icl;
            const double B = cr * irr * fB * kr;
            const double p_t = ics * (cr * A - cl * B);
            const double fU = 1.0 - urho * cl * icr2;
            const double Pu = A + cr * irr * fU * kr;
            const double Pp = cr * A - cl * B;
            const double Ye = erx - pr * icr2 * prx + 2.0 * (er - e) * icr * crx;
            // ---- reverse
            b_vr += Wb.ut;
            double b_urho = 0, b_icr2 = 0, b_A = 0, b_B = 0, b_r3 = 0, b_irho = 0, b_den = 0;
            // v_t = -um rho irr vrx
            b_um -= rho * irr * vrx * b_v_t;
            b_rho -= um * irr * vrx * b_v_t;
            b_irr -= um * rho * vrx * b_v_t;
            b_vrx -= um * rho * irr * b_v_t;
            // e_t = -urho irr Ye + pm icr2 p_t
            b_urho -= irr * Ye * b_e_t;
            b_irr -= urho * Ye * b_e_t;
            const double b_Ye = -urho * irr * b_e_t;
            b_pm += icr2 * p_t * b_e_t;
            b_icr2 += pm * p_t * b_e_t;
            b_p_t += pm * icr2 * b_e_t;
            b_erx += b_Ye;
            b_pr -= icr2 * prx * b_Ye;
            b_icr2 -= pr * prx * b_Ye;
            b_prx -= pr * icr2 * b_Ye;
            b_er += 2.0 * icr * crx * b_Ye;
            b_e -= 2.0 * icr * crx * b_Ye;
            b_icr += 2.0 * (er - e) * crx * b_Ye;
            b_crx += 2.0 * (er - e) * icr * b_Ye;
            // u_t = ics (A + cr irr fU kr)
            b_ics += Pu * b_u_t;
            const double b_Pu = ics * b_u_t;
            b_A += b_Pu;
            b_cr += irr * fU * kr * b_Pu;
            b_irr += cr * fU * kr * b_Pu;
            const double b_fU = cr * irr * kr * b_Pu;
            b_kr += cr * irr * fU * b_Pu;
            b_urho -= cl * icr2 * b_fU;
            b_cl -= urho * icr2 * b_fU;
            b_icr2 -= urho * cl * b_fU;
            // p_t = ics (cr A - cl B)
            b_ics += Pp * b_p_t;
            const double b_Pp = ics * b_p_t;
            b_cr += A * b_Pp;
            b_A += cr * b_Pp;
            b_cl -= B * b_Pp;
            b_B -= cl * b_Pp;
            // B = cr irr fB kr, fB = 1 - urho icl
            b_cr += irr * fB * kr * b_B;
            b_irr += cr * fB * kr * b_B;
            const double b_fB = cr * irr * kr * b_B;
            b_kr += cr * irr * fB * b_B;
            b_urho -= icl * b_fB;
            b_icl -= urho * b_fB;
            // A = cl irl fA kl, fA = 1 + urho icr
            b_cl += irl * fA * kl * b_A;
            b_irl += cl * fA * kl * b_A;
            const double b_fA = cl * irl * kl * b_A;
            b_kl += cl * irl * fA * b_A;
            b_urho += icr * b_fA;
            b_icr += urho * b_fA;
            // urho = um rho
            b_um += rho * b_urho;
            b_rho += um * b_urho;
            // rho_t = X1 t1 + X2 Dp
            const double b_X2 = Dp * b_rho_t;
            b_Dp += X2 * b_rho_t;
            b_r3 += icr2 * icr * Z * b_X2;
            b_icr2 += r3 * icr * Z * b_X2;
            b_icr += r3 * icr2 * Z * b_X2;
            const double b_Z = r3 * icr2 * icr * b_X2;
            b_um += b_Z;
            b_cr += irho * b_Z;
            b_irho += cr * b_Z;
            const double b_X1 = t1 * b_rho_t;
            const double b_t1 = X1 * b_rho_t;
            b_um -= r3 * irr * icr2 * b_X1;
            b_r3 -= um * irr * icr2 * b_X1;
            b_irr -= um * r3 * icr2 * b_X1;
            b_icr2 -= um * r3 * irr * b_X1;
            // t1 = cr^2 irr^2 rrx - cr urx + 1.5 (um - ur) crx
            b_cr += (2.0 * cr * irr * irr * rrx - urx) * b_t1;
            b_irr += 2.0 * cr * cr * irr * rrx * b_t1;
            b_rrx += cr * cr * irr * irr * b_t1;
            b_urx -= cr * b_t1;
            b_um += 1.5 * crx * b_t1;
            b_ur -= 1.5 * crx * b_t1;
            b_crx += 1.5 * (um - ur) * b_t1;
            // r3 = rho^3
            b_rho += 3.0 * rho * rho * b_r3;
            // e = er + dpm spm icr2 / 2
            b_er += b_e;
            const double b_dpm = spm * 0.5 * icr2 * b_e;
            const double b_spm = dpm * 0.5 * icr2 * b_e;
            b_icr2 += dpm * spm * 0.5 * b_e;
            b_pm += b_dpm + b_spm;
            b_pr += -b_dpm + b_spm;
            b_icr += 2.0 * icr * b_icr2;
            // irho = den icr ; rho = cr / den
            b_den += icr * b_irho;
            b_icr += den * b_irho;
            b_cr += iden * b_rho;
            b_den -= iden * iden * cr * b_rho;
            // den = ur - um + cr irr
            b_ur += b_den;
            b_um -= b_den;
            b_cr += irr * b_den;
            b_irr += cr * b_den;
        } else {
            // ---- forward of the left sub-case
            const double den = um - ul + cl * irl;
            const double iden = rcp(den);
            const double rho = cl * iden;
            const double irho = den * icl;
            const double icl2 = icl * icl;
            const double dpm = pm - pl, spm = pm + pl;
            const double e = el + dpm * spm * 0.5 * icl2;
            const double r3 = rho * rho * rho;
            const double t1 = cl * cl * irl * irl * rlx + cl * ulx + 1.5 * (um - ul) * clx;
            const double X1 = -um * r3 * irl * icl2;
            const double Z = um - cl * irho;
            const double X2 = -r3 * icl2 * icl * Z;
            const double urho = um * rho;
            const double fA = 1.0 + urho * icr;
            const double A = cl * irl * fA * kl;
            const double fB = 1.0 - urho * icl;
            const double B = cr * irr * fB * kr;
            const double p_t = ics * (cr * A - cl * B);
            const double fU = 1.0 + urho * cr * icl2;
            const double Pu = cl * irl * fU * kl + B;
            const double Pp = cr * A - cl * B;
            const double Ye = elx - pl * icl2 * plx + 2.0 * (el - e) * icl * clx;
            // ---- reverse
            b_vl += Wb.ut;
            double b_urho = 0, b_icl2 = 0, b_A = 0, b_B = 0, b_r3 = 0, b_irho = 0, b_den = 0;
            b_um -= rho * irl * vlx * b_v_t;
            b_rho -= um * irl * vlx * b_v_t;
            b_irl -= um * rho * vlx * b_v_t;
            b_vlx -= um * rho * irl * b_v_t;
            b_urho -= irl * Ye * b_e_t;
            b_irl -= urho * Ye * b_e_t;
            const double b_Ye = -urho * irl * b_e_t;
            b_pm += icl2 * p_t * b_e_t;
            b_icl2 += pm * p_t * b_e_t;
            b_p_t += pm * icl2 * b_e_t;
            b_elx += b_Ye;
            b_pl -= icl2 * plx * b_Ye;
            b_icl2 -= pl * plx * b_Ye;
            b_plx -= pl * icl2 * b_Ye;
            b_el += 2.0 * icl * clx * b_Ye;
            b_e -= 2.0 * icl * clx * b_Ye;
            b_icl += 2.0 * (el - e) * clx * b_Ye;
            b_clx += 2.0 * (el - e) * icl * b_Ye;
            // u_t = ics (cl irl fU kl + B)
            b_ics += Pu * b_u_t;
            const double b_Pu = ics * b_u_t;
            b_cl += irl * fU * kl * b_Pu;
            b_irl += cl * fU * kl * b_Pu;
            const double b_fU = cl * irl * kl * b_Pu;
            b_kl += cl * irl * fU * b_Pu;
            b_B += b_Pu;
            b_urho += cr * icl2 * b_fU;
            b_cr += urho * icl2 * b_fU;
            b_icl2 += urho * cr * b_fU;
            // p_t = ics (cr A - cl B)
            b_ics += Pp * b_p_t;
            const double b_Pp = ics * b_p_t;
            b_cr += A * b_Pp;
            b_A += cr * b_Pp;
            b_cl -= B * b_Pp;
            b_B -= cl * b_Pp;
            b_cr += irr * fB * kr * b_B;
            b_irr += cr * fB * kr * b_B;
            const double b_fB = cr * irr * kr * b_B;
            b_kr += cr * irr * fB * b_B;
            b_urho -= icl * b_fB;
            b_icl -= urho * b_fB;
            b_cl += irl * fA * kl * b_A;
            b_irl += cl * fA * kl * b_A;
            const double b_fA = cl * irl * kl * b_A;
            b_kl += cl * irl * fA * b_A;
            b_urho += icr * b_fA;
            b_icr += urho * b_fA;
            b_um += rho * b_urho;
            b_rho += um * b_urho;
            // rho_t = X1 t1 + X2 Dp, X2 = -r3 icl2 icl Z, Z = um - cl irho
            const double b_X2 = Dp * b_rho_t;
            b_Dp += X2 * b_rho_t;
            b_r3 -= icl2 * icl * Z * b_X2;
            b_icl2 -= r3 * icl * Z * b_X2;
            b_icl -= r3 * icl2 * Z * b_X2;
            const double b_Z = -r3 * icl2 * icl * b_X2;
            b_um += b_Z;
            b_cl -= irho * b_Z;
            b_irho -= cl * b_Z;
            const double b_X1 = t1 * b_rho_t;
            const double b_t1 = X1 * b_rho_t;
            b_um -= r3 * irl * icl2 * b_X1;
            b_r3 -= um * irl * icl2 * b_X1;
            b_irl -= um * r3 * icl2 * b_X1;
            b_icl2 -= um * r3 * irl * b_X1;
            // t1 = cl^2 irl^2 rlx + cl ulx + 1.5 (um - ul) clx
            b_cl += (2.0 * cl * irl * irl * rlx + ulx) * b_t1;
            b_irl += 2.0 * cl * cl * irl * rlx * b_t1;
            b_rlx += cl * cl * irl * irl * b_t1;
            b_ulx += cl * b_t1;
            b_um += 1.5 * clx * b_t1;
            b_ul -= 1.5 * clx * b_t1;
            b_clx += 1.5 * (um - ul) * b_t1;
            b_rho += 3.0 * rho * rho * b_r3;
            b_el += b_e;
            const double b_dpm = spm * 0.5 * icl2 * b_e;
            const double b_spm = dpm * 0.5 * icl2 * b_e;
            b_icl2 += dpm * spm * 0.5 * b_e;
            b_pm += b_dpm + b_spm;
            b_pl += -b_dpm + b_spm;
            b_icl += 2.0 * icl * b_icl2;
            b_den += icl * b_irho;
            b_icl += den * b_irho;
            b_cl += iden * b_rho;
            b_den -= iden * iden * cl * b_rho;
            // den = um - ul + cl irl
            b_um += b_den;
            b_ul -= b_den;
            b_cl += irl * b_den;
            b_irl += cl * b_den;
        }
        // ---- reverse of the common star part
        // Dp = ND ics / 2, ND = cr dl - cl dr
        const double b_ND = 0.5 * ics * b_Dp;
        b_ics += ND * 0.5 * b_Dp;
        b_cr += dl * b_ND;
        const double b_dl = cr * b_ND;
        b_cl -= dr * b_ND;
        const double b_dr = -cl * b_ND;
        // dr = cr irr Qr, Qr = 2 cr urx + ur crx - 2 prx
        b_cr += irr * Qr * b_dr;
        b_irr += cr * Qr * b_dr;
        const double b_Qr = cr * irr * b_dr;
        b_cr += 2.0 * urx * b_Qr;
        b_urx += 2.0 * cr * b_Qr;
        b_ur += crx * b_Qr;
        b_crx += ur * b_Qr;
        b_prx -= 2.0 * b_Qr;
        // dl = cl irl Ql, Ql = -2 cl ulx - ul clx - 2 plx
        b_cl += irl * Ql * b_dl;
        b_irl += cl * Ql * b_dl;
        const double b_Ql = cl * irl * b_dl;
        b_cl -= 2.0 * ulx * b_Ql;
        b_ulx -= 2.0 * cl * b_Ql;
        b_ul -= clx * b_Ql;
        b_clx -= ul * b_Ql;
        b_plx -= 2.0 * b_Ql;
        // kr = -prx + cr urx + ur crx / 2
        b_prx -= b_kr;
        b_cr += urx * b_kr;
        b_urx += cr * b_kr;
        b_ur += 0.5 * crx * b_kr;
        b_crx += 0.5 * ur * b_kr;
        // kl = -plx - cl ulx - ul clx / 2
        b_plx -= b_kl;
        b_cl -= ulx * b_kl;
        b_ulx -= cl * b_kl;
        b_ul -= 0.5 * clx * b_kl;
        b_clx -= 0.5 * ul * b_kl;
        // pm = Np ics
        const double b_Np = ics * b_pm;
        b_ics += Np * b_pm;
        b_du += cl * cr * b_Np;
        b_cl += (du * cr + pr) * b_Np;
        b_cr += (du * cl + pl) * b_Np;
        b_pr += cl * b_Np;
        b_pl += cr * b_Np;
        // um = Nu ics
        const double b_Nu = ics * b_um;
        b_ics += Nu * b_um;
        b_ul += cl * b_Nu;
        b_cl += ul * b_Nu;
        b_ur += cr * b_Nu;
        b_cr += ur * b_Nu;
        b_pl += b_Nu;
        b_pr -= b_Nu;
        // ics = 1 / (cl + cr)
        const double b_S = -ics * ics * b_ics;
        b_cl += b_S;
        b_cr += b_S;
    }
    // ---- relaxation speeds
    double b_Sl = 0, b_Sr = 0, b_iSl = 0, b_iSr = 0;
    if (P) {
        b_cr -= icr * icr * b_icr;
        b_Sr += b_cr;
        b_rr += kappa * mr * b_cr;
        const double b_mr = rr * kappa * b_cr;
        const double b_xr = xr > 0.0 ? b_mr : 0.0;
        b_pl += icl * b_xr;
        b_pr -= icl * b_xr;
        b_icl += (pl - pr) * b_xr;
        b_du += b_xr;
        b_cl -= icl * icl * b_icl;
        b_Sl += b_cl;
        b_rl += kappa * ml * b_cl;
        const double b_ml = rl * kappa * b_cl;
        const double b_xl = xl > 0.0 ? b_ml : 0.0;
        b_pr += iSr * b_xl;
        b_pl -= iSr * b_xl;
        b_iSr += (pr - pl) * b_xl;
        b_du += b_xl;
    } else {
        b_cl -= icl * icl * b_icl;
        b_Sl += b_cl;
        b_rl += kappa * ml * b_cl;
        const double b_ml = rl * kappa * b_cl;
        const double b_xl = xl > 0.0 ? b_ml : 0.0;
        b_pr += icr * b_xl;
        b_pl -= icr * b_xl;
        b_icr += (pr - pl) * b_xl;
        b_du += b_xl;
        b_cr -= icr * icr * b_icr;
        b_Sr += b_cr;
        b_rr += kappa * mr * b_cr;
        const double b_mr = rr * kappa * b_cr;
        const double b_xr = xr > 0.0 ? b_mr : 0.0;
        b_pl += iSl * b_xr;
        b_pr -= iSl * b_xr;
        b_iSl += (pl - pr) * b_xr;
        b_du += b_xr;
    }
    // du = ul - ur
    b_ul += b_du;
    b_ur -= b_du;
    // clx = gamma/2 iSl sxl, sxl = pl rlx + rl plx (same on the right)
    b_iSl += 0.5 * gamma * sxl * b_clx;
    {
        const double b_s = 0.5 * gamma * iSl * b_clx;
        b_pl += rlx * b_s;
        b_rlx += pl * b_s;
        b_rl += plx * b_s;
        b_plx += rl * b_s;
    }
    b_iSr += 0.5 * gamma * sxr * b_crx;
    {
        const double b_s = 0.5 * gamma * iSr * b_crx;
        b_pr += rrx * b_s;
        b_rrx += pr * b_s;
        b_rr += prx * b_s;
        b_prx += rr * b_s;
    }
    // iS = 1/S, S = sqrt(gamma p rho)
    b_Sl -= iSl * iSl * b_iSl;
    b_Sr -= iSr * iSr * b_iSr;
    b_pl += 0.5 * gamma * rl * iSl * b_Sl;
    b_rl += 0.5 * gamma * pl * iSl * b_Sl;
    b_pr += 0.5 * gamma * rr * iSr * b_Sr;
    b_rr += 0.5 * gamma * pr * iSr * b_Sr;
    // e_x = z irl g1i, z = p_x - p irl rho_x
    {
        const double b_z = irl * g1i * b_elx;
        b_irl += zl * g1i * b_elx;
        b_plx += b_z;
        b_pl -= irl * rlx * b_z;
        b_irl -= pl * rlx * b_z;
        b_rlx -= pl * irl * b_z;
    }
    {
        const double b_z = irr * g1i * b_erx;
        b_irr += zr * g1i * b_erx;
        b_prx += b_z;
        b_pr -= irr * rrx * b_z;
        b_irr -= pr * rrx * b_z;
        b_rrx -= pr * irr * b_z;
    }
    // e = p irl g1i
    b_pl += irl * g1i * b_el;
    b_irl += pl * g1i * b_el;
    b_pr += irr * g1i * b_er;
    b_irr += pr * g1i * b_er;
    // irl = 1/rl
    b_rl -= irl * irl * b_irl;
    b_rr -= irr * irr * b_irr;
    flb = Q{b_rl, b_pl, b_ul, b_vl};
    frb = Q{b_rr, b_pr, b_ur, b_vr};
    slb = Q{b_rlx, b_plx, b_ulx, b_vlx};
    srb = Q{b_rrx, b_prx, b_urx, b_vrx};
}

// ---------------------------------------------------------------- update
// Adjoint of q' = update<false>(q, FL, FR, lambda) given qb' (adjoint of q').
// Returns Ub = d(q')/dU^T qb' (adjoint of the conserved update), adds the
// adjoint of q, and returns dq'/dlambda . qb' through lamb.
FVM_HD Q prim_adj(const Q &qo, const Q &qob, const Gas &G) {
    // q' = (c0, g1 (c2 - (c1^2 + c3^2)/(2 c0)), c1/c0, c3/c0) at q' = qo
    const double ir = rcp(qo.r);
    const double un = qo.un, ut = qo.ut;
    Q Ub;
    Ub.r = qob.r + qob.p * G.g1 * 0.5 * (un * un + ut * ut) - (qob.un * un + qob.ut * ut) * ir;
    Ub.p = -qob.p * G.g1 * un + qob.un * ir;
    Ub.un = qob.p * G.g1;
    Ub.ut = -qob.p * G.g1 * ut + qob.ut * ir;
    return Ub;
}

// Adjoint of cons(q) = (rho, rho un, p/(gamma-1) + rho (un^2 + ut^2)/2, rho ut).
FVM_HD Q cons_adj(const Q &q, const Q &Ub, const Gas &G) {
    return Q{Ub.r + q.un * Ub.p + 0.5 * (q.un * q.un + q.ut * q.ut) * Ub.un + q.ut * Ub.ut,
             G.g1i * Ub.un, q.r * Ub.p + q.r * q.un * Ub.un, q.r * Ub.ut + q.r * q.ut * Ub.un};
}

// Gradient of the wave speed |(u, v)| + sqrt(gamma p / rho) (zero velocity
// part at rest).
FVM_HD Q speed_grad(double rho, double p, double u, double v, const Gas &G) {
    const double a = sqrt(G.gamma * p / rho);
    const double vel = sqrt(u * u + v * v);
    const double ia = 0.5 / a;
    Q g;
    g.r = -G.gamma * p / (rho * rho) * ia;
    g.p = G.gamma / rho * ia;
    g.un = vel > 0.0 ? u / vel : 0.0;
    g.ut = vel > 0.0 ? v / vel : 0.0;
    return g;  // (d/drho, d/dp, d/du, d/dv)
}

}  // namespace fvm2d
