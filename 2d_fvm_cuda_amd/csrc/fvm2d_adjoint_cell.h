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
        // ---- forward: both sub-cases through side K (u* <= 0: K = r, s = -1;
        // else K = l, s = +1) in the closed form of grp<false>: with
        // P = k_l/rho_l, R = k_r/rho_r and M = (c_l P + c_r R)/(c_l + c_r),
        // Dpi/Dt = c_l c_r (P - R)/(c_l + c_r), p_t = Dpi/Dt + u* rho* M,
        // u_t = M + u* rho* Dpi/Dt / c_K^2, rho_t = (-u* (c_K/rho_K) t1 +
        // (c_K/rho_K - s u_K) Dpi/Dt) / den^3 with den = c_K / rho*
        const double ics = rcp(cl + cr);
        const double Nu = ul * cl + ur * cr + pl - pr;
        const double um = Nu * ics;
        const double Np = du * cl * cr + cl * pr + cr * pl;
        const double pm = Np * ics;
        const double kl = -plx - cl * ulx - 0.5 * ul * clx;
        const double kr = -prx + cr * urx + 0.5 * ur * crx;
        const double Pl = kl * irl, Rr = kr * irr;
        const double D = Pl - Rr;
        const double clcr = cl * cr;
        const double Dp = clcr * ics * D;
        const double Mn = cl * Pl + cr * Rr;
        const double M = Mn * ics;
        const bool right = um <= 0.0;
        const double s = right ? -1.0 : 1.0;
        const double cK = right ? cr : cl, icK = right ? icr : icl, irK = right ? irr : irl;
        const double uK = right ? ur : ul, pK = right ? pr : pl;
        const double eKx = right ? erx : elx, cKx = right ? crx : clx;
        const double rxK = right ? rrx : rlx, uxK = right ? urx : ulx;
        const double pxK = right ? prx : plx, vxK = right ? vrx : vlx;
        const double cirK = cK * irK;
        const double den = s * (um - uK) + cirK;
        const double iden = rcp(den);
        const double rho = cK * iden;
        const double icK2 = icK * icK;
        const double dpm = pm - pK, spm = pm + pK;
        const double D2 = dpm * spm * icK2;  // (p*^2 - p_K^2) / c_K^2 = 2 (e* - e_K)
        const double t1 = cirK * cirK * rxK + s * cK * uxK + 1.5 * (um - uK) * cKx;
        const double iden3 = iden * iden * iden;
        const double umc = um * cirK;
        const double G1 = cirK - s * uK;
        const double Hn = -umc * t1 + G1 * Dp;  // rho_t = iden3 Hn
        const double urho = um * rho;
        const double qK = umc * iden;  // u* rho* / rho_K
        const double p_t = Dp + urho * M;
        const double Ye = -D2 * icK * cKx - pK * icK2 * pxK + eKx;  // e_t = -qK Ye + pm icK2 p_t
        // ---- reverse
        double b_rho = Wb.r, b_pm = Wb.p, b_um = Wb.un;
        const double b_e = Wb.e, b_vK = Wb.ut;
        const double b_rho_t = Wtb.r, b_u_t = Wtb.un, b_v_t = Wtb.ut, b_e_t = Wtb.e;
        double b_p_t = Wtb.p;
        // v_t = -qK vxK
        double b_qK = -vxK * b_v_t;
        const double b_vxK = -qK * b_v_t;
        // e_t = -qK Ye + pm icK2 p_t
        b_qK -= Ye * b_e_t;
        const double b_Ye = -qK * b_e_t;
        b_pm += icK2 * p_t * b_e_t;
        double b_icK2 = pm * p_t * b_e_t;
        b_p_t += pm * icK2 * b_e_t;
        // Ye = -D2 icK cKx - pK icK2 pxK + eKx
        double b_D2 = -icK * cKx * b_Ye;
        double b_icK = -D2 * cKx * b_Ye;
        double b_cKx = -D2 * icK * b_Ye;
        double b_pK = -icK2 * pxK * b_Ye;
        b_icK2 -= pK * pxK * b_Ye;
        const double b_pxK = -pK * icK2 * b_Ye;
        const double b_eKx = b_Ye;
        // u_t = M + urho icK2 Dp
        double b_M = b_u_t;
        double b_urho = icK2 * Dp * b_u_t;
        b_icK2 += urho * Dp * b_u_t;
        double b_Dp = urho * icK2 * b_u_t;
        // p_t = Dp + urho M
        b_Dp += b_p_t;
        b_urho += M * b_p_t;
        b_M += urho * b_p_t;
        // qK = umc iden
        double b_umc = iden * b_qK;
        double b_iden = umc * b_qK;
        // urho = um rho
        b_um += rho * b_urho;
        b_rho += um * b_urho;
        // rho_t = iden3 Hn, Hn = -umc t1 + G1 Dp
        const double b_iden3 = Hn * b_rho_t;
        const double b_Hn = iden3 * b_rho_t;
        b_umc -= t1 * b_Hn;
        const double b_t1 = -umc * b_Hn;
        b_Dp += G1 * b_Hn;
        // G1 = cirK - s uK
        double b_cirK = Dp * b_Hn;
        double b_uK = -s * Dp * b_Hn;
        // umc = um cirK
        b_um += cirK * b_umc;
        b_cirK += um * b_umc;
        // iden3 = iden^3
        b_iden += 3.0 * iden * iden * b_iden3;
        // t1 = cirK^2 rxK + s cK uxK + 1.5 (um - uK) cKx
        b_cirK += 2.0 * cirK * rxK * b_t1;
        const double b_rxK = cirK * cirK * b_t1;
        double b_cK = s * uxK * b_t1;
        const double b_uxK = s * cK * b_t1;
        b_um += 1.5 * cKx * b_t1;
        b_uK -= 1.5 * cKx * b_t1;
        b_cKx += 1.5 * (um - uK) * b_t1;
        // e = eK + D2 / 2
        b_D2 += 0.5 * b_e;
        // D2 = dpm spm icK2
        b_pm += (spm + dpm) * icK2 * b_D2;
        b_pK += (dpm - spm) * icK2 * b_D2;
        b_icK2 += dpm * spm * b_D2;
        // icK2 = icK^2
        b_icK += 2.0 * icK * b_icK2;
        // rho = cK iden
        b_cK += iden * b_rho;
        b_iden += cK * b_rho;
        // iden = 1 / den; den = s (um - uK) + cirK
        const double b_den = -iden * iden * b_iden;
        b_um += s * b_den;
        b_uK -= s * b_den;
        b_cirK += b_den;
        // cirK = cK irK
        b_cK += irK * b_cirK;
        const double b_irK = cK * b_cirK;
        // side K back to the left or right quantities
        auto side = [right](double b, double &to_l, double &to_r) {
            to_l += right ? 0.0 : b;
            to_r += right ? b : 0.0;
        };
        side(b_cK, b_cl, b_cr);
        side(b_icK, b_icl, b_icr);
        side(b_irK, b_irl, b_irr);
        side(b_uK, b_ul, b_ur);
        side(b_pK, b_pl, b_pr);
        side(b_e, b_el, b_er);
        side(b_eKx, b_elx, b_erx);
        side(b_cKx, b_clx, b_crx);
        side(b_rxK, b_rlx, b_rrx);
        side(b_uxK, b_ulx, b_urx);
        side(b_pxK, b_plx, b_prx);
        side(b_vxK, b_vlx, b_vrx);
        side(b_vK, b_vl, b_vr);
        // M = Mn ics, Mn = cl Pl + cr Rr
        const double b_Mn = ics * b_M;
        double b_ics = Mn * b_M;
        b_cl += Pl * b_Mn;
        double b_Pl = cl * b_Mn;
        b_cr += Rr * b_Mn;
        double b_Rr = cr * b_Mn;
        // Dp = clcr ics D, D = Pl - Rr
        const double b_clcr = ics * D * b_Dp;
        b_ics += clcr * D * b_Dp;
        const double b_D = clcr * ics * b_Dp;
        b_cl += cr * b_clcr;
        b_cr += cl * b_clcr;
        b_Pl += b_D;
        b_Rr -= b_D;
        // Pl = kl irl, Rr = kr irr
        const double b_kl = irl * b_Pl;
        b_irl += kl * b_Pl;
        const double b_kr = irr * b_Rr;
        b_irr += kr * b_Rr;
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
