// carbchem_dev.h -- seawater carbonate chemistry on the device (FP64).
//
// Restates gem_carbchem.f90: sub_calc_carbconst (59-268, Mehrbach set),
// the salinity-derived totals (834-890), sub_calc_carb (326-546: the
// fixed-point [H+] iteration H <- sqrt(H1*H2) with the stop test of line 482)
// and sub_calc_carb_RF0 (553-669).  Operand order follows the reference; the
// transcendental functions are the device library's (within 1 ulp of glibc).
#pragma once
#include "cgocn_dev.h"

namespace cgo {

enum { CC_K = 0, CC_K1, CC_K2, CC_KB, CC_KW, CC_KSI, CC_KHF, CC_KHSO4, CC_KP1, CC_KP2, CC_KP3, CC_KCAL, CC_KARG,
       CC_QCO2, CC_QO2, CC_KH2S, CC_KNH4, NCC };

__device__ __forceinline__ double corr_p(double TC, double P, double rRT, double a, double b, double c, double d,
                                         double e) {
    return (-(a + b * TC + c * TC * TC) + (5.0E-4 * (d + e * TC)) * P) * P * rRT;
}
__device__ __forceinline__ double salt_Ca(double S) { double v = 0.01028 * S / 35.0; return v < NULLSMALL ? NULLSMALL : v; }
__device__ __forceinline__ double salt_B(double S) { double v = 0.000416 * S / 35.0; return v < NULLSMALL ? NULLSMALL : v; }
__device__ __forceinline__ double salt_F(double S) { double v = 0.00007 * S / 35.0; return v < NULLSMALL ? NULLSMALL : v; }
__device__ __forceinline__ double salt_SO4(double S) { double v = 0.02824 * S / 35.0; return v < NULLSMALL ? NULLSMALL : v; }

__device__ inline void calc_carbconst(double D, double T, double S, double* cc) {
    const double CONST_R = 83.145;
    double lT = T, lS = S;
    if (lT < (ZEROC + 2.0)) lT = ZEROC + 2.0;
    if (lT > (ZEROC + 35.0)) lT = ZEROC + 35.0;
    if (lS < 26.0) lS = 26.0;
    if (lS > 43.0) lS = 43.0;
    const double P = D / 10.0;
    const double S05 = sqrt(lS), S15 = pow(lS, 1.5), S20 = lS * lS;
    const double Tln = log(lT), Tlog = log10(lT), rT = 1.0 / lT, Tr100 = lT / 100.0;
    const double TC = lT - ZEROC, rRT = 1.0 / (CONST_R * lT);
    double I = lS > NULLSMALL ? 19.924 * lS / (1000.0 - 1.005 * lS) : NULLSMALL;
    const double I05 = sqrt(I), I15 = pow(I, 1.5), I20 = I * I;
    double Cl = S / 1.80655;
    if (Cl < NULLSMALL) Cl = NULLSMALL;
    const double ION = Cl > NULLSMALL ? 0.00147 + 0.03592 * Cl + 0.000068 * Cl * Cl : NULLSMALL;
    const double ION05 = sqrt(ION);
    const double m2c = log(1 - 0.001005 * lS);
    const double SO4tot = salt_SO4(lS), Ftot = salt_F(lS);
    const double lnkHSO4 = 141.328 - 4276.1 * rT - 23.093 * Tln + (324.57 - 13856.0 * rT - 47.986 * Tln) * I05 +
                           (-771.54 + 35474.0 * rT + 114.723 * Tln) * I - 2698.0 * rT * I15 + 1776.0 * rT * I20;
    const double lnkHF = 1590.2 / lT - 12.641 + 1.525 * ION05;
    cc[CC_KHSO4] = exp(lnkHSO4 + m2c);
    const double f2t = log(1.0 + SO4tot / cc[CC_KHSO4]);
    cc[CC_KHF] = exp(lnkHF + m2c + f2t);
    const double f2s = log(1.0 + SO4tot / cc[CC_KHSO4] + Ftot / cc[CC_KHF]);
    const double t2s = -f2t + f2s;
    const double lnk1 = log(pow(10.0, -(3670.7 * rT - 62.008 + 9.7944 * Tln - 0.0118 * lS + 0.000116 * S20)));
    const double lnk2 = log(pow(10.0, -(1394.7 * rT + 4.777 - 0.0184 * lS + 0.000118 * S20)));
    cc[CC_K1] = exp(lnk1 + corr_p(TC, P, rRT, -2.550E+1, +1.271E-1, +0.000E+0, -3.080E+0, +8.770E-2));
    cc[CC_K2] = exp(lnk2 + corr_p(TC, P, rRT, -1.582E+1, -2.190E-2, +0.000E+0, +1.130E+0, -1.475E-1));
    cc[CC_K] = cc[CC_K1] / cc[CC_K2];
    const double lnkB = 148.0248 + 137.194 * S05 + 1.62247 * lS +
                        (-8966.90 - 2890.51 * S05 - 77.942 * lS + 1.726 * S15 - 0.0993 * S20) * rT +
                        (-24.4344 - 25.085 * S05 - 0.2474 * lS) * Tln + 0.053105 * S05 * lT;
    cc[CC_KB] = exp(lnkB + m2c + t2s + corr_p(TC, P, rRT, -2.948E+1, +1.622E-1, +2.608E-3, -2.840E+0, +0.000E+0));
    const double lnkW = 148.9802 - 13847.26 * rT - 23.6521 * Tln + (-5.977 + 118.67 * rT + 1.0495 * Tln) * S05 -
                        0.01615 * lS;
    cc[CC_KW] = exp(lnkW + corr_p(TC, P, rRT, -2.002E+1, +1.119E-1, -1.409E-3, -5.130E+0, +7.940E-2));
    const double lnkSi = 117.40 - 8904.2 * rT - 19.334 * Tln + (3.5913 - 458.79 * rT) * I05 +
                         (-1.5998 + 188.74 * rT) * I + (0.07871 - 12.1652 * rT) * I * I;
    cc[CC_KSI] = exp(lnkSi + m2c + corr_p(TC, P, rRT, -2.948E+1, +1.622E-1, +2.608E-3, -2.840E+0, +0.000E+0));
    cc[CC_KHF] = exp(lnkHF + m2c + f2s + corr_p(TC, P, rRT, -9.780E+0, -9.000E-3, -9.420E-4, -3.910E+0, +5.400E-2));
    cc[CC_KHSO4] = exp(lnkHSO4 + m2c + f2s +
                       corr_p(TC, P, rRT, -1.803E+1, +4.660E-2, +3.160E-4, -4.530E+0, +9.000E-2));
    const double lnkP1 = 115.54 - 4576.752 / lT - 18.453 * Tln + (0.69171 - 106.736 / lT) * S05 +
                         (-0.01844 - 0.65643 / lT) * lS;
    cc[CC_KP1] = exp(lnkP1 + corr_p(TC, P, rRT, -1.451E+1, +1.211E-1, -3.210E-4, -2.670E+0, +4.270E-2));
    const double lnkP2 = 172.1033 - 8814.715 / lT - 27.927 * Tln + (1.3566 - 160.340 / lT) * S05 +
                         (-0.05778 + 0.37335 / lT) * lS;
    cc[CC_KP2] = exp(lnkP2 + corr_p(TC, P, rRT, -2.312E+1, +1.758E-1, -2.647E-3, -5.150E+0, +9.000E-2));
    const double lnkP3 = -18.126 - 3070.75 / lT + (2.81197 + 17.27039 / lT) * S05 + (-0.09984 - 44.99486 / lT) * lS;
    cc[CC_KP3] = exp(lnkP3 + corr_p(TC, P, rRT, -2.657E+1, +2.020E-1, -3.042E-3, -4.080E+0, +7.140E-2));
    const double lnkH2S = 225.838 - 13275.3 * rT - 34.6435 * Tln + 0.3449 * S05 - 0.0274 * lS;
    cc[CC_KH2S] = exp(lnkH2S + t2s + corr_p(TC, P, rRT, -1.107E+1, +9.000E-3, -9.420E-4, +2.890E+0, +5.400E-2));
    const double lnkNH4 = -6285.33 * rT + 0.0001635 * lT - 0.25444 + (0.46532 - 123.7184 * rT) * S05 +
                          (-0.01992 + 3.17556 * rT) * lS;
    cc[CC_KNH4] = exp(lnkNH4 + corr_p(TC, P, rRT, -2.643E+0, +8.890E-1, -9.050E-3, -5.030E+0, +8.140E-2));
    const double logkcal = -171.9065 - 0.077993 * lT + 2839.319 * rT + 71.595 * Tlog +
                           (-0.77712 + 0.0028426 * lT + 178.34 * rT) * S05 - 0.07711 * lS + 0.0041249 * S15;
    const double logkarg = -171.945 - 0.077993 * lT + 2903.293 * rT + 71.595 * Tlog +
                           (-0.068393 + 0.0017276 * lT + 88.135 * rT) * S05 - 0.10018 * lS + 0.0059415 * S15;
    cc[CC_KCAL] = exp(corr_p(TC, P, rRT, -4.876E+1, +5.304E-1, +0.000E+0, -1.176E+1, +3.692E-1)) * pow(10.0, logkcal);
    cc[CC_KARG] = exp(corr_p(TC, P, rRT, -4.596E+1, +5.304E-1, +0.000E+0, -1.176E+1, +3.692E-1)) * pow(10.0, logkarg);
    const double lnQCO2 = -60.2409 + 93.4517 * (100 * rT) + 23.3585 * log(Tr100) +
                          lS * (0.023517 - 0.023656 * (Tr100) + 0.0047036 * (Tr100 * Tr100));
    const double lnQO2 = -173.9894 + 255.5907 * (100.0 * rT) + 146.4813 * log(Tr100) - 22.2040 * (Tr100) +
                         lS * (-0.037362 + 0.016504 * (Tr100) - 0.0020564 * (Tr100 * Tr100)) - log(1.0E6) -
                         log(0.20946);
    cc[CC_QCO2] = exp(lnQCO2);
    cc[CC_QO2] = exp(lnQO2);
}

struct Spec { double CO2, CO3, HCO3, H1, H2; };

// one speciation sweep at [H] (gem_carbchem.f90:355-427); SiO2, H2S and NH4
// are zero for the species the engine carries
__device__ __forceinline__ void carb_sweep(double H, double DIC, double ALK, double PO4t, double Bt, double SO4t,
                                           double Ft, const double* cc, Spec& o) {
    const double H2 = H * H, H3 = H * H2;
    const double OH = cc[CC_KW] / H;
    const double H4BO4 = Bt / (1.0 + H / cc[CC_KB]);
    const double H3SiO4 = 0.0 / (1.0 + H / cc[CC_KSI]);
    const double HSO4 = SO4t / (1.0 + cc[CC_KHSO4] / H);
    const double HF = Ft / (1.0 + cc[CC_KHF] / H);
    const double P1 = cc[CC_KP1], P2 = cc[CC_KP2], P3 = cc[CC_KP3];
    const double H3PO4 = PO4t / (1.0 + P1 / H + (P1 * P2) / H2 + (P1 * P2 * P3) / H3);
    const double HPO4 = PO4t / (1.0 + H / P2 + H2 / (P1 * P2) + P3 / H);
    const double PO4 = PO4t / (1.0 + H / P3 + H2 / (P2 * P3) + H3 / (P1 * P2 * P3));
    const double A = ALK - H4BO4 - OH - HPO4 - 2.0 * PO4 - H3SiO4 - 0.0 - 0.0 + H + HSO4 + HF + H3PO4;
    const double K = cc[CC_K];
    const double t = 4.0 * A + DIC * K - A * K;
    const double zed = sqrt(t * t + 4.0 * (K - 4.0) * (A * A));
    o.HCO3 = (DIC * K - zed) / (K - 4.0);
    o.CO3 = (A * K - DIC * K - 4.0 * A + zed) / (2.0 * (K - 4.0));
    o.CO2 = DIC - A + (A * K - DIC * K - 4.0 * A + zed) / (2.0 * (K - 4.0));
    o.H1 = cc[CC_K1] * o.CO2 / o.HCO3;
    o.H2 = cc[CC_K2] * o.HCO3 / o.CO3;
}

// 1/x to ~1 ulp: the hardware reciprocal and two Newton steps (the FP32 form's
// members; FP64 members divide exactly as the reference)
__device__ __forceinline__ double rcp_nr(double x) {
    double r = __builtin_amdgcn_rcp(x);
    r = fma(fma(-x, r, 1.0), r, r);
    r = fma(fma(-x, r, 1.0), r, r);
    return r;
}

// The loop-invariant reciprocals of one solve (FP32 form)
struct CarbRcp {
    double rKB, rP2, rP3, rP1P2, rP2P3, rP1P2P3, P1, P1P2, P1P2P3, rK4, r2K4;
};
__device__ __forceinline__ void carb_rcp(const double* cc, CarbRcp& r) {
    const double P1 = cc[CC_KP1], P2 = cc[CC_KP2], P3 = cc[CC_KP3];
    r.P1 = P1;
    r.P1P2 = P1 * P2;
    r.P1P2P3 = P1 * P2 * P3;
    r.rKB = 1.0 / cc[CC_KB];
    r.rP2 = 1.0 / P2;
    r.rP3 = 1.0 / P3;
    r.rP1P2 = 1.0 / r.P1P2;
    r.rP2P3 = 1.0 / (P2 * P3);
    r.rP1P2P3 = 1.0 / r.P1P2P3;
    r.rK4 = 1.0 / (cc[CC_K] - 4.0);
    r.r2K4 = 1.0 / (2.0 * (cc[CC_K] - 4.0));
}

// the speciation sweep of carb_sweep with the divides replaced by products
// with reciprocals (FP32 form); the same fixed point to rounding
__device__ __forceinline__ void carb_sweep_fast(double H, double DIC, double ALK, double PO4t, double Bt, double SO4t,
                                                double Ft, const double* cc, const CarbRcp& rc, Spec& o) {
    const double rH = rcp_nr(H);
    const double H2 = H * H, H3 = H * H2, rH2 = rH * rH, rH3 = rH2 * rH;
    const double OH = cc[CC_KW] * rH;
    const double H4BO4 = Bt * rcp_nr(1.0 + H * rc.rKB);
    const double HSO4 = SO4t * rcp_nr(1.0 + cc[CC_KHSO4] * rH);
    const double HF = Ft * rcp_nr(1.0 + cc[CC_KHF] * rH);
    const double H3PO4 = PO4t * rcp_nr(1.0 + rc.P1 * rH + rc.P1P2 * rH2 + rc.P1P2P3 * rH3);
    const double HPO4 = PO4t * rcp_nr(1.0 + H * rc.rP2 + H2 * rc.rP1P2 + cc[CC_KP3] * rH);
    const double PO4 = PO4t * rcp_nr(1.0 + H * rc.rP3 + H2 * rc.rP2P3 + H3 * rc.rP1P2P3);
    const double A = ALK - H4BO4 - OH - HPO4 - 2.0 * PO4 + H + HSO4 + HF + H3PO4;
    const double K = cc[CC_K];
    const double t = 4.0 * A + DIC * K - A * K;
    const double zed = sqrt(t * t + 4.0 * (K - 4.0) * (A * A));
    o.HCO3 = (DIC * K - zed) * rc.rK4;
    o.CO3 = (A * K - DIC * K - 4.0 * A + zed) * rc.r2K4;
    o.CO2 = DIC - A + o.CO3;
    o.H1 = cc[CC_K1] * o.CO2 * rcp_nr(o.HCO3);
    o.H2 = cc[CC_K2] * o.HCO3 * rcp_nr(o.CO3);
}

// sub_calc_carb: >0 iterations, -1 negative [H] estimate, -2 iteration limit
// (ctrl_carbchem_fail: carb is left unchanged, no random re-seed)
__device__ inline int calc_carb(double DIC, double ALK, double Ca, double PO4, double Bt, double SO4t, double Ft,
                                const double* cc, AS_GLOBAL double* carb, size_t cs, double tol, int maxit) {
    int n = 1;
    double H = carb[IC_H * cs];
    Spec o;
    for (;;) {
        const double Hold = H;
        carb_sweep(H, DIC, ALK, PO4, Bt, SO4t, Ft, cc, o);
        if ((o.H1 < NULLSMALL) || (o.H2 < NULLSMALL)) return -1;
        H = sqrt(o.H1 * o.H2);
        if (fabs(1.0 - H / Hold) < (1.0E-8 / H) * tol) {
            carb[IC_CONC_CO2 * cs] = o.CO2;
            carb[IC_CONC_CO3 * cs] = o.CO3;
            carb[IC_CONC_HCO3 * cs] = o.HCO3;
            carb[IC_FUG_CO2 * cs] = o.CO2 / cc[CC_QCO2];
            carb[IC_OHM_CAL * cs] = Ca * o.CO3 / cc[CC_KCAL];
            carb[IC_OHM_ARG * cs] = Ca * o.CO3 / cc[CC_KARG];
            carb[IC_H * cs] = H;
            carb[IC_DCO3_CAL * cs] = o.CO3 - cc[CC_KCAL] * 1.0 / Ca;
            carb[IC_DCO3_ARG * cs] = o.CO3 - cc[CC_KARG] * 1.0 / Ca;
            return n;
        }
        n = n + 1;
        if (n > maxit) return -2;
    }
}

__device__ inline int calc_carb_RF0(double DIC, double ALK, double PO4, double Bt, double SO4t, double Ft,
                                    const double* cc, AS_GLOBAL double* carb, size_t cs, double tol, int maxit) {
    int n = 1;
    double H = carb[IC_H * cs];
    const double DICr = DIC + 1.0e-6;
    Spec o;
    for (;;) {
        const double Hold = H;
        carb_sweep(H, DICr, ALK, PO4, Bt, SO4t, Ft, cc, o);
        H = sqrt(o.H1 * o.H2);
        if (fabs(1.0 - H / Hold) < tol) {
            carb[IC_RF0 * cs] = (o.CO2 / carb[IC_CONC_CO2 * cs] - 1.0) / (DICr / DIC - 1.0);
            return n;
        }
        n = n + 1;
        if ((o.H1 < NULLSMALL) || (o.H2 < NULLSMALL) || (n > maxit)) {
            carb[IC_RF0 * cs] = 0.0;
            return -1;
        }
    }
}

// The FP32 form's members: the same fixed point H = sqrt(H1(H) H2(H)) with
// the reference's stop tests, accelerated by Steffensen's method (two plain
// sweeps, then Aitken's delta-squared extrapolation).  It converges in a few
// sweeps where the plain iteration takes ~25 per step; the root agrees with
// the plain iteration's to the stop tolerance.  As in the reference, the
// stored speciation is that of the last sweep and [H] its output.
__device__ inline double aitken(double h0, double h1, double h2) {
    const double d = (h2 - h1) - (h1 - h0);
    const double hn = d != 0.0 ? h0 - (h1 - h0) * (h1 - h0) / d : h2;
    return (hn > 0.0 && hn < 2.0 * h2 + h0) ? hn : h2;  // a bad extrapolation keeps the plain step
}

__device__ inline int calc_carb_fast(double DIC, double ALK, double Ca, double PO4, double Bt, double SO4t, double Ft,
                                     const double* cc, AS_GLOBAL double* carb, size_t cs, double tol, int maxit) {
    int n = 0;
    double H = carb[IC_H * cs];
    Spec o;
    CarbRcp rc;
    carb_rcp(cc, rc);
    double h[3];
    for (;;) {
        h[0] = H;
        bool done = false;
        for (int j = 0; j < 2 && !done; j++) {
            const double Hin = H;
            carb_sweep_fast(Hin, DIC, ALK, PO4, Bt, SO4t, Ft, cc, rc, o);
            n = n + 1;
            if ((o.H1 < NULLSMALL) || (o.H2 < NULLSMALL)) return -1;
            H = sqrt(o.H1 * o.H2);
            h[j + 1] = H;
            done = fabs(1.0 - H / Hin) < (1.0E-8 / H) * tol;
        }
        if (done) break;
        if (n >= maxit) return -2;
        H = aitken(h[0], h[1], h[2]);
    }
    carb[IC_CONC_CO2 * cs] = o.CO2;
    carb[IC_CONC_CO3 * cs] = o.CO3;
    carb[IC_CONC_HCO3 * cs] = o.HCO3;
    carb[IC_FUG_CO2 * cs] = o.CO2 / cc[CC_QCO2];
    carb[IC_OHM_CAL * cs] = Ca * o.CO3 / cc[CC_KCAL];
    carb[IC_OHM_ARG * cs] = Ca * o.CO3 / cc[CC_KARG];
    carb[IC_H * cs] = H;
    carb[IC_DCO3_CAL * cs] = o.CO3 - cc[CC_KCAL] * 1.0 / Ca;
    carb[IC_DCO3_ARG * cs] = o.CO3 - cc[CC_KARG] * 1.0 / Ca;
    return n;
}

__device__ inline int calc_carb_RF0_fast(double DIC, double ALK, double PO4, double Bt, double SO4t, double Ft,
                                         const double* cc, AS_GLOBAL double* carb, size_t cs, double tol, int maxit) {
    int n = 0;
    double H = carb[IC_H * cs];
    const double DICr = DIC + 1.0e-6;
    Spec o;
    CarbRcp rc;
    carb_rcp(cc, rc);
    double h[3];
    for (;;) {
        h[0] = H;
        bool done = false;
        for (int j = 0; j < 2 && !done; j++) {
            const double Hin = H;
            carb_sweep_fast(Hin, DICr, ALK, PO4, Bt, SO4t, Ft, cc, rc, o);
            n = n + 1;
            H = sqrt(o.H1 * o.H2);
            h[j + 1] = H;
            done = fabs(1.0 - H / Hin) < tol;
            if (!done && ((o.H1 < NULLSMALL) || (o.H2 < NULLSMALL) || (n > maxit))) {
                carb[IC_RF0 * cs] = 0.0;
                return -1;
            }
        }
        if (done) break;
        H = aitken(h[0], h[1], h[2]);
    }
    carb[IC_RF0 * cs] = (o.CO2 / carb[IC_CONC_CO2 * cs] - 1.0) / (DICr / DIC - 1.0);
    return n;
}

}  // namespace cgo
