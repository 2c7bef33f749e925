/*
 * device_math.h -- per-cell closures evaluated inside the HIP kernels:
 * Tillotson and tabulated EOS, strength/damage/tensile registry, remap
 * helpers.  Each follows the reference function named beside it, keeping its
 * operand order (kernels are compiled with -ffp-contract=off).
 */
#pragma once
#include "salec_internal.h"

namespace salec {

#define TOLVOF 1.0e-6
#define TOLRHO 1.0e-4
#define TOLQPRE 1e-4
#define DTF 0.2
#define CISSAMK1 1.0
#define BETA_JC 0.57735026918962576451

/* src/eos/eos_state.c:7-54 */
__device__ __forceinline__ double dWind(double x, double l, double r) { return x < l ? l : (x > r ? r : x); }
__device__ __forceinline__ double dMax(double a, double b) { return a > b ? a : b; }
__device__ __forceinline__ double dMin(double a, double b) { return a < b ? a : b; }
__device__ __forceinline__ double dSqrtTrunc(double x) { return x < 1.0e-2 ? -sqrt(fabs(x)) : sqrt(x); }

/* src/geometry/linear.c:188-194 */
__device__ __forceinline__ double second_inv(const double *t)
{
    double tXX = t[0], tYY = t[2], tXY = t[1], tTH = t[3];
    return ((tXX - tYY) * (tXX - tYY) + (tYY - tTH) * (tYY - tTH) + (tTH - tXX) * (tTH - tXX)) / 6.0 + tXY * tXY;
}

/* src/geometry/linear.c:196-218 */
__device__ __forceinline__ int bisearch(double x, const double *xl, int n)
{
    int left = 0, right = n - 1;
    if (x < xl[left]) return left;
    if (x >= xl[right]) return right - 1;
    while (left + 1 < right) {
        int mid = (left + right) / 2;
        if (xl[mid] > x) right = mid; else left = mid;
    }
    return left;
}

/* bisearch with a hint (AxisHint): the same interval index, found from a
 * log-spaced start instead of ~log2(n) dependent loads.  For x in
 * [x[1], x[n-1]) the walk below ends at the unique i with x[i] <= x < x[i+1],
 * which is what bisearch converges to on a strictly increasing axis; the
 * edges and NaN take bisearch's own branches. */
__device__ __forceinline__ int axis_index(double x, const double *xl, int n, const AxisHint &h)
{
    if (h.nb == 0 || x != x) return bisearch(x, xl, n);
    if (x < xl[1]) return 0;
    if (x >= xl[n - 1]) return n - 2;
    double u = (log(x) - h.l1) * h.inv;
    int b = u < 0. ? 0 : (u >= (double)(h.nb - 1) ? h.nb - 1 : (int)u);
    int i = h.start[b];
    while (i > 0 && xl[i] > x) --i;
    while (i < n - 2 && xl[i + 1] <= x) ++i;
    return i;
}

/* TillPres with TillCold/TillHot/TillTran, src/eos/eos_state.c:322-438 */
__device__ inline void till_pres(const DevMat &t, double E, double R, double *P, double *Cs)
{
    double Eta = R / t.TLRho0, Mu = Eta - 1.0;
    double Phi0 = E / (t.TLE0 * Eta * Eta), Phi1 = 1.0 / (Phi0 + 1.0);
    double B = Eta < 1.0 ? t.TLA : t.TLB;
    if ((R >= t.TLRho0) || (E < t.TLEiv)) {
        *P = (t.TLa + t.TLb * Phi1) * E * R + t.TLA * Mu + B * Mu * Mu;
        double c1 = (1.0 / t.TLRho0) * (t.TLA + 2.0 * B * Mu) + E * (t.TLa + t.TLb * Phi1 * (3.0 - 2.0 * Phi1));
        double c2 = (t.TLa + t.TLb * Phi1 * Phi1) * (*P) / R;
        *Cs = dSqrtTrunc(c1 + c2);
    } else {
        double Xi = t.TLRho0 / R - 1.0;
        double A2 = 0.0; if (t.TLAlpha * Xi * Xi < 100.0) A2 = exp(-1.0 * t.TLAlpha * Xi * Xi);
        double B1 = 0.0; if (t.TLBeta * Xi < 100.0) B1 = exp(-1.0 * t.TLBeta * Xi);
        double Ph = t.TLa * R * E + (t.TLb * E * R * Phi1 + t.TLA * Mu * B1) * A2;
        double ch1 = t.TLa * E + t.TLb * E * Phi1 * (3.0 + 2.0 * t.TLAlpha * Xi / Eta - 2.0 * Phi1) * A2
                   + t.TLA / t.TLRho0 * (1.0 - (Xi / Eta) * (t.TLBeta + 2.0 * t.TLAlpha * Xi)) * A2 * B1;
        if (E > t.TLEcv) {
            *P = Ph;
            double c2 = (t.TLa + t.TLb * Phi1 * Phi1 * A2) * (*P) / R;
            *Cs = dSqrtTrunc(ch1 + c2);
        } else {
            double ET = t.TLEcv - t.TLEiv, dEr = E - t.TLEiv, dEl = t.TLEcv - E;
            double Pc = (t.TLa + t.TLb * Phi1) * E * R + t.TLA * Mu + B * Mu * Mu;
            *P = (dEl * Pc + dEr * Ph) / ET;
            double cc1 = (1.0 / t.TLRho0) * (t.TLA + 2.0 * B * Mu) + E * (t.TLa + t.TLb * Phi1 * (3.0 - 2.0 * Phi1));
            double cc2 = (t.TLa + t.TLb * Phi1 * Phi1) * (*P) / R;
            double ch2 = (t.TLa + t.TLb * Phi1 * Phi1 * A2) * (*P) / R;
            double c2c = cc1 + cc2, c2h = ch1 + ch2;
            double s1 = (dEl * c2c + dEr * c2h) / ET;
            double s2 = (Ph - Pc) / ET * (*P) / (R * R);
            *Cs = dSqrtTrunc(s1 + s2);
        }
    }
    if (*P <= 1.0e-2) *Cs = dSqrtTrunc(t.TLA / t.TLRho0);
}

/* TillTemp, src/eos/eos_state.c:519-526 */
__device__ inline double till_temp(const DevMat &t, double E, double R)
{
    int DL = (int)(dWind(floor(R / t.till_stepden), 0, t.till_nden - 2));
    double Dxl = DL + 1.0 - R / t.till_stepden;
    double Dxr = 1.0 - Dxl;
    double rE = Dxl * t.till_yeng[DL] + Dxr * t.till_yeng[DL + 1];
    return dMax((E - rE) / t.TLCv + t.TLTref, 0.0);
}

#define DTAB(t, i, j, k) ((t).tab_data[((long)(i) * (t).tab_nden + (j)) * 3 + (k)])

/* aneossieden_calculate_state, src/eos/aneos_sieden.c:264-289 */
__device__ inline void sd_state(const DevMat &t, double den, double eng, double *tem, double *pre, double *csd)
{
    double tDen = den, tEng = eng / den;
    int DI = axis_index(tDen, t.tab_yden, t.tab_nden, t.hden), DJ = DI + 1;
    int TI = axis_index(tEng, t.tab_xtem, t.tab_ntem, t.heng), TJ = TI + 1;
    double Dxr = (tDen - t.tab_yden[DI]) / (t.tab_yden[DJ] - t.tab_yden[DI]);
    double Dxl = dWind(1.0 - Dxr, 0.0, 1.0);
    double Exr = (tEng - t.tab_xtem[TI]) / (t.tab_xtem[TJ] - t.tab_xtem[TI]);
    double Exl = dWind(1.0 - Exr, 0., 1.0);
    double a0 = DTAB(t, TI, DI, 0), a1 = DTAB(t, TI, DI, 1), a2 = DTAB(t, TI, DI, 2);
    double b0 = DTAB(t, TI, DJ, 0), b1 = DTAB(t, TI, DJ, 1), b2 = DTAB(t, TI, DJ, 2);
    double c0 = DTAB(t, TJ, DI, 0), c1 = DTAB(t, TJ, DI, 1), c2 = DTAB(t, TJ, DI, 2);
    double d0 = DTAB(t, TJ, DJ, 0), d1 = DTAB(t, TJ, DJ, 1), d2 = DTAB(t, TJ, DJ, 2);
    *tem = (a0 * Dxl + b0 * Dxr) * Exl + (c0 * Dxl + d0 * Dxr) * Exr;
    *pre = (a1 * Dxl + b1 * Dxr) * Exl + (c1 * Dxl + d1 * Dxr) * Exr;
    *csd = (a2 * Dxl + b2 * Dxr) * Exl + (c2 * Dxl + d2 * Dxr) * Exr;
}

/* aneos_calculate_state (temperature-density table), src/eos/eos_state.c:793-841 */
__device__ inline void aneos_state(const DevMat &t, double den, double eng, double *tem, double *pre, double *csd)
{
    int DI = 0, DJ = t.tab_nden - 1;
    while (DI + 1 < DJ) { int DM = (DI + DJ) / 2; if (t.tab_yden[DM] > den) DJ = DM; else DI = DM; }
    double Dxr = (den - t.tab_yden[DI]) / (t.tab_yden[DJ] - t.tab_yden[DI]);
    Dxr = dWind(Dxr, 0.0, 1.0);
    double Dxl = 1.0 - Dxr;
    int TI = 0;
    double IRE = DTAB(t, TI, DI, 0) * Dxl + DTAB(t, TI, DJ, 0) * Dxr;
    int TJ = t.tab_ntem - 1;
    double JRE = DTAB(t, TJ, DI, 0) * Dxl + DTAB(t, TJ, DJ, 0) * Dxr;
    while (TI + 1 < TJ) {
        int TM = (TI + TJ) / 2;
        double MRE = DTAB(t, TM, DI, 0) * Dxl + DTAB(t, TM, DJ, 0) * Dxr;
        if (MRE > eng) { TJ = TM; JRE = MRE; } else { TI = TM; IRE = MRE; }
    }
    double Txr = (eng - IRE) / (JRE - IRE);
    Txr = dWind(Txr, 0.0, 1.0);
    double Txl = 1.0 - Txr;
    *tem = t.tab_xtem[TI] * Txl + t.tab_xtem[TJ] * Txr;
    *pre = (DTAB(t, TI, DI, 1) * Dxl + DTAB(t, TI, DJ, 1) * Dxr) * Txl + (DTAB(t, TJ, DI, 1) * Dxl + DTAB(t, TJ, DJ, 1) * Dxr) * Txr;
    *csd = (DTAB(t, TI, DI, 2) * Dxl + DTAB(t, TI, DJ, 2) * Dxr) * Txl + (DTAB(t, TJ, DI, 2) * Dxl + DTAB(t, TJ, DJ, 2) * Dxr) * Txr;
}

/* calculate_state dispatch (src/eos/eos_state.c:727-745, 844-850) followed by
 * calculate_state_correct with CAP_MINPRESSURE (src/eos/eos_state.c:747-791) */
__device__ inline void eos_state_capped(const DevMat &m, double den, double eng, double plim, double pty,
                                        double *tem, double *pre, double *csd)
{
    double tt = 0., tp = 0., tc = 0.;
    double d = den * pty, e = eng * pty;
    if (m.eos_type == SALEC_EOS_TILLOTSON) {
        double E = e / d;
        tt = till_temp(m, E, d);
        till_pres(m, E, d, &tp, &tc);
    } else if (m.eos_type == SALEC_EOS_ANEOS_SD) {
        sd_state(m, d, e, &tt, &tp, &tc);
    } else if (m.eos_type == SALEC_EOS_ANEOS) {
        aneos_state(m, d, e, &tt, &tp, &tc);
    }
    tp = tp / pty;
    if (pty > 1.0) tc *= (1.0 + (m.chi - 1.0) * fmax(0.0, fmin(1.0, (pty - 1.0) / (m.alpha0 - 1.0))));
    if (tp < plim) tp = dMax(tp, plim);
    *tem = tt; *pre = tp; *csd = tc;
}

/* src/strength.c:505-575 */
__device__ __forceinline__ double Ylundborg(double p, double y0, double fric, double ylim) { return y0 + p * fric / (1 + p * fric / (ylim - y0)); }
__device__ __forceinline__ double Ydrucker(double p, double y0, double fric, double ylim) { return dMin(y0 + fric * p, ylim); }
__device__ inline double SimonMelt(double pre, const DevMat &s)
{
    double st = s.Tmelt0 * pow(s.Asimon * pre + 1.0, s.Csimon);
    if (s.PolyLiq <= 0 && s.PolySol <= 0) return st;
    if (s.PolySol > 0) { double p = pre * 1.0e-9; st = s.Tsolc1 + s.Tsolc2 * p + s.Tsolc3 * p * p; }
    double p = pre * 1.0e-9;
    double lt = 0.0;
    lt = s.Tlidc1 + s.Tlidc2 * p + s.Tlidc3 * p * p;
    return dMax(st, lt);
}
__device__ __forceinline__ double OhnakaSoft(double tem, const DevMat &s, double Tm) { return tanh(((Tm + s.Tdelta) / tem - 1.0) * s.Tfrac); }
/* src/strength.c:724-751 */
__device__ inline double IceMelt(double p)
{
    const double p0 = 0.0, p1 = 0.2e9, p2 = 0.34e9, t0 = 273.0, t1 = 252.0, t2 = 257.0;
    const double p3 = 0.61e9, p4 = 2.17e9, t3 = 274.0, t4 = 354.8;
    if (p <= p1) return t0 + (t1 - t0) * (p - p0) / (p1 - p0);
    else if (p > p1 && p <= p2) return t1 + (t2 - t1) * (p - p1) / (p2 - p1);
    else if (p > p2 && p <= p3) return t2 + (t3 - t2) * (p - p2) / (p3 - p2);
    else if (p > p3 && p <= p4) return t3 + (t4 - t3) * (p - p3) / (p4 - p3);
    else return t4 * pow(((p - p4) / 1.253e9) + 1, 1.0 / 3.0);
}

/* strength registry (str_func_list), src/strength.c:379-503, 697-791 */
__device__ inline double strength_fn(int fn, double p0, double p1, double p2, double p3, double p4, const DevMat &s)
{
    switch (fn) {
    case SALEC_FN_SIMPLEROCK: {
        double Yint = Ylundborg(p0, s.Yint0, s.Yfricint, s.Ylimint);
        double Ydam = Ydrucker(p0, s.Ydam0, s.Yfricdam, s.Ylimdam);
        Ydam = dMin(Yint, Ydam);
        return Ydam * p1 + Yint * (1.0 - p1);
    }
    case SALEC_FN_SIMPLESHEAR: case SALEC_FN_IVANOV:
        return 1.0 / dMax(s.IvanA, s.IvanB * (p0 - s.IvanC));
    case SALEC_FN_JOHNSONCOOK1: case SALEC_FN_JOHNSONCOOK2: {
        double PS = dWind(p2, 0.0, s.JCMaxPlastic);
        double Y1 = s.JCA + s.JCB * pow(PS, s.JCN), Y2 = 1.0;
        if (p3 > 1.0 && s.JCC > 0.0) Y2 = 1.0 + s.JCC * log(p3);
        if (fn == SALEC_FN_JOHNSONCOOK2) {
            double Tm = SimonMelt(p0, s);
            double f = (p4 - s.JCTREF) / (Tm - s.JCTREF);
            f = dWind(f, 0.0, 1.0);
            double soft = dWind(1.0 - pow(f, s.JCM), 0.0, 1.0);
            return Y1 * Y2 * soft * BETA_JC;
        }
        return Y1 * Y2 * BETA_JC;
    }
    case SALEC_FN_ROCKSTRENGTHT: {
        double Yint = Ylundborg(p0, s.Yint0, s.Yfricint, s.Ylimint);
        double Ydam = Ydrucker(p0, s.Ydam0, s.Yfricdam, s.Ylimdam);
        Ydam = dMin(Yint, Ydam);
        double mt = SimonMelt(p0, s);
        double sf = OhnakaSoft(p4, s, mt);
        sf = dWind(sf, 0.0, 1.0);
        double ymed = (Ydam * p1 + Yint * (1.0 - p1)) * sf;
        double yvis = p3 * s.melt_viscosity * (1.0 - sf);
        return dMax(ymed, yvis);
    }
    case SALEC_FN_ICESTRENGTH: {
        double Yint = Ylundborg(p0, s.Yint0, s.Yfricint, s.Ylimint);
        double Ydam = Ylundborg(p0, s.Ydam0, s.Yfricdam, s.Ylimdam);
        Ydam = dMin(Yint, Ydam);
        double sf = OhnakaSoft(p4, s, IceMelt(p0));
        sf = dWind(sf, 0.0, 1.0);
        return (Ydam * p1 + Yint * (1.0 - p1)) * sf;
    }
    case SALEC_FN_COLLINS: {
        double inc = 0.;
        if (p0 < s.Pbd) inc = dMax(0.01, 0.01 + 0.04 * p0 / s.Pbd);
        else if (p0 < s.Pbp) inc = 0.05 + 0.05 * (p0 - s.Pbd) / (s.Pbp - s.Pbd);
        else inc = 0.1 + 0.5 * (p0 - s.Pbp) / s.Pbp;
        return 1.0 / inc;
    }
    case SALEC_FN_SIMPLETENSILE: {
        double mt = SimonMelt(p0, s);
        double sf = OhnakaSoft(p4, s, mt);
        sf = dWind(sf, 0.0, 1.0);
        return s.Yten0 * (1.0 - p1) * sf;
    }
    case SALEC_FN_VONMISES1: {
        double Yint = s.Yint0, Ydam = dMin(Yint, s.Ydam0);
        return Ydam * p1 + Yint * (1.0 - p1);
    }
    case SALEC_FN_VONMISES2: {
        double Yint = s.Yint0, Ydam = dMin(Yint, s.Ydam0);
        double sf = OhnakaSoft(p4, s, SimonMelt(p0, s));
        sf = dWind(sf, 0.0, 1.0);
        return (Ydam * p1 + Yint * (1.0 - p1)) * sf;
    }
    default:
        return 0.0;
    }
}

/* get_flux_volcyl, src/advect.c:1347-1354 */
__device__ __forceinline__ double flux_volcyl(double xa, double ya, double xb, double yb,
                                              double dxa, double dya, double dxb, double dyb)
{
    double x1 = -1.0 * dxa, y1 = -1.0 * dya, x2 = -1.0 * dxb, y2 = -1.0 * dyb;
    double vol = x1 * (ya - yb - y2) + x2 * (ya - yb + y1) + (y1 + y2) * (xb - xa);
    return vol * fabs(2.0 * xa + 2.0 * xb + x1 + x2) / 8.0;
}

/* cut_length, src/advect.c:1356-1408 (arguments already ordered a >= b) */
__device__ inline double cut_length(double vf, double a, double b)
{
    if (a < b) { double t = a; a = b; b = t; }
    if (1.0e-5 * a > b) {
        double d = 0.0;
        if (vf <= 0.0) d = 0.0; else if (vf <= 1.0) d = vf * a; else d = a;
        return d + b;
    }
    double d, crf1 = 0.5 * b / a;
    if (vf <= crf1) d = sqrt(2.0 * a * b * vf);
    else if (vf <= 1.0 - crf1) d = 0.5 * b + a * vf;
    else if (vf <= 1.0) d = a + b - sqrt(2.0 * a * b * (1.0 - vf));
    else d = a + b;
    return d;
}

/* geometric_flux_ratio, src/advect.c:1312-1330 */
__device__ __forceinline__ double geo_ratio(double p1, double p2)
{
    double a = p1, b = p2;
    if (b > a) { a = p2; b = p1; }
    if (b >= 0.0) return 1.0;
    else if (a <= 0.0) return 0.0;
    else return a / (a + fabs(b));
}

/* vofMSTACS, src/advect.c:1259-1274 */
__device__ __forceinline__ double mstacs(double nv, double gf, double cd)
{
    double cds, sto;
    if (cd < 1.0 / 3.0) cds = dMin(nv / cd, 1.0); else cds = dMin(3.0 * nv, 1.0);
    if (nv < 1.0 / 5.0) sto = 3.0 * nv;
    else if (nv < 1.0 / 2.0) sto = (1.0 + nv) / 2.0;
    else sto = dMin((6.0 * nv + 3.0) / 8.0, 1.0);
    return gf * cds + (1.0 - gf) * sto;
}

/* algebraic_flux_ratio, src/advect.c:1226-1252 */
__device__ inline double alg_ratio(double U, double D, double A, double cd, double bn0, double bn1, double g0, double g1)
{
    double tmp = A - U;
    double nD = (D - U) / tmp;
    if (fabs(tmp) < TOLVOF) return D;
    if (nD < TOLVOF || nD > 1.0 - TOLVOF) return D;
    double z = 0;
    z += g0 * bn0; z += g1 * bn1;
    double l = 0.;
    l += g0 * g0; l += g1 * g1;
    double ct = z / sqrt(l);
    ct = ct * ct;
    double gf = dMin(CISSAMK1 * ct * ct, 1.0);
    double nf = mstacs(nD, gf, cd);
    return U + nf * (A - U);
}

/* vec_norm with the NaN -> 1.0 quirk, src/geometry/linear.c:116-131 */
__device__ __forceinline__ void vnorm2(double &a, double &b)
{
    double l = 0.;
    l += a * a; l += b * b;
    l = sqrt(l);
    a /= l; if (isnan(a)) a = 1.0;
    b /= l; if (isnan(b)) b = 1.0;
}

/* Interpolate2d with IpB = {lb, rb, ru, lu}, src/geometry/interpolate.c:36-44 */
__device__ __forceinline__ double interp2d(double x0, double x1, double x2, double x3, double r, double s)
{
    double v = 0;
    v += 0.25 * (1 - r) * (1 - s) * x0;
    v += 0.25 * (1 + r) * (1 - s) * x1;
    v += 0.25 * (1 + r) * (1 + s) * x2;
    v += 0.25 * (1 - r) * (1 + s) * x3;
    return v;
}

} // namespace salec
