// k_bgc.hip -- BIOGEM column step and biogem_tracercoupling on the device.
//
// Built with -ffp-contract=off: operand order follows the reference so the
// FP64 results track the CPU restatement to the last bits (only the device
// exp/log/pow differ from glibc, within 1 ulp).
//
// k_bgc_col   one thread per (member, wet column), everything it carries
//             down the column held in registers: the surface carbonate
//             chemistry (biogem.f90:1026-1084), then one march from
//             the surface to k1 that applies, level by level, DOM
//             remineralisation (biogem_box.f90:2287-2406), particulate
//             remineralisation (2412-2877), the sediment return of the
//             previous step's rain (biogem.f90:873-942) and PO4 uptake
//             (biogem_box.f90:346-1515).  Every cell is read and written once.
// k_couple_sums    one block per (member, 64 wet columns): the column partial
//             sums of the state after transport (biogem.f90:1984-1997,
//             2016-2025).
// k_couple_reduce  one block per member: the global sums, in the reference's
//             column order (biogem.f90:1927-2031), staged through LDS.
// k_couple_apply   one block per (member, 64 wet columns): the rescale and
//             re-normalisation of every tracer, T, S, masses and
//             particulates (2001-2061).
#include "carbchem_dev.h"
#include "tstep_impl.h"

#include <algorithm>

namespace cgo {

// conv_sed_ocn for the carried species (gem_util.f90:79-98,
// biogem_data.f90:751-789); columns B_DIC, B_PO4, B_O2, B_ALK
struct Conv {
    double c[NSED][4];
};

__device__ __forceinline__ void make_conv(const BgcParams& p, Conv& cv) {
    for (int s = 0; s < NSED; s++)
        for (int q = 0; q < 4; q++) cv.c[s][q] = 0.0;
    const double cO2POP = -4.0 / 2.0;
    const double cO2PON = 0.0;
    const double cO2POC = p.red_POP_PO2 / p.red_POP_POC - cO2POP / p.red_POP_POC - cO2PON * p.red_POP_PON / p.red_POP_POC;
    cv.c[S_POC][B_DIC] = 1.0;
    cv.c[S_POC][B_O2] = cO2POC;
    cv.c[S_POP][B_PO4] = 1.0;
    cv.c[S_POP][B_O2] = cO2POP;
    cv.c[S_POP][B_ALK] = p.red_PON_ALK * p.red_POP_PON;
    cv.c[S_CACO3][B_DIC] = 1.0;
    cv.c[S_CACO3][B_ALK] = 2.0;
}

// sub_box_remin_redfield (biogem_box.f90:2100-2203) with O2 the only oxidant
__device__ __forceinline__ double redfield(const BgcParams& p, double O2) {
    if (p.slot[B_O2] < 0) return 0.0;
    double lk = 0.0;
    if (O2 < 0.0) O2 = 0.0;
    double kO2 = O2 / (O2 + p.remin_c0_O2);
    lk = lk + p.remin_k_O2 * kO2;
    if (lk < NULLSMALL) lk = 1.0;
    if (O2 < NULLSMALL) kO2 = 1.0;
    return p.remin_k_O2 * kO2 / lk;
}

// ---------------------------------------------------------------- air-sea
// area of a cell in row J (biogem_data.f90:1156, atchem_data.f90:222)
__device__ __forceinline__ double cell_area(const AS_GLOBAL double* sv, int maxi, int J) {
    return 2.0 * CONST_PI * (CONST_REARTH * CONST_REARTH) * (1.0 / maxi) * (sv[J] - sv[J - 1]);
}

// fun_calc_solconst (gem_carbchem.f90:1367-1426) on the clamped T and S:
// CO2 in mol kg-1 atm-1 (Weiss 1974), O2 from its Bunsen coefficient over
// rho const_V; coefficient tables of gem_data.f90:68-141
__device__ __forceinline__ double solconst(double b0, double b1, double b2, double b3, double b4, double b5,
                                           bool bunsen, double T, double S, double rho) {
    const double lT = T < ZEROC + 2.0 ? ZEROC + 2.0 : (T > (ZEROC + 35.0) ? ZEROC + 35.0 : T);
    const double lS = S < 26.0 ? 26.0 : (S > 43.0 ? 43.0 : S);
    const double rT = 1.0 / lT, Tr100 = lT / 100.0;
    const double x = exp(b0 + b1 * (100 * rT) + b2 * log(Tr100) + lS * (b3 + b4 * (Tr100) + b5 * (Tr100 * Tr100)));
    return bunsen ? x / (rho * CONST_V) : x;
}

// sub_calc_pv (biogem_box.f90:81-117): Wanninkhof (1992) piston velocity, m yr-1
__device__ __forceinline__ double piston(double c0, double c1, double c2, double c3, double T, double u, double a) {
    double TC = T - ZEROC;
    if (TC < 0.0) TC = 0.0;
    if (TC > 30.0) TC = 30.0;
    const double TC2 = TC * TC, TC3 = TC2 * TC;
    const double Sc = c0 - c1 * TC + c2 * TC2 - c3 * TC3;
    return CONV_CM_M * CONV_YR_HR * a * (u * u) * pow(Sc * 1.515E-3, -0.5);
}

// fun_calc_ocnatm_flux for one gas (biogem_box.f90:123-299), ocean -> air
// positive, mol yr-1: one step moves at most the surface layer's
// disequilibrium (par_airsea_r_dflux_deqm_max = 1, biogem_lib.f90:708)
__device__ __forceinline__ double ocnatm_flux(double pv, double A, double rho, double dD, double ocnv, double atmv,
                                              double buff, double dtyr) {
    if (ocnv < NULLSMALL) ocnv = 0.0;
    if (atmv < NULLSMALL) atmv = 0.0;
    const double lA = (1.0 - 0.0) * A;
    double fo = pv * lA * rho * ocnv;
    double fa = pv * lA * rho * atmv;
    const double deqm = dD * A * rho * buff * fabs(atmv - ocnv);
    const double dflux = dtyr * fabs(fo - fa);
    if (deqm > NULLSMALL) {
        const double r = dflux / deqm;
        if (r > 1.00) {
            fo = (1.00 / r) * fo;
            fa = (1.00 / r) * fa;
        }
    }
    return fo - fa;
}

// remineralised sediment -> ocean tracers: v[q] += (rf * conv[s][q]) * x[s]
__device__ __forceinline__ void add_remin(const BgcParams& p, const Conv& cv, double rf, const double* x, double* v) {
#pragma unroll
    for (int s = 0; s < NSED; s++) {
        if (s == S_CACO3 && !p.caco3) continue;
#pragma unroll
        for (int q = 0; q < 4; q++)
            if (p.slot[q] >= 0 && cv.c[s][q] != 0.0) v[q] = v[q] + (rf * cv.c[s][q]) * x[s];
    }
}

// one level of a sinking particle stream (the kk body of the settling loop of
// sub_box_remin_part, biogem_box.f90:2555-2693): t holds the flux leaving
// level kk+1 and becomes the flux leaving kk; prem is what dissolves in kk
__device__ __forceinline__ void stream_step(const BgcParams& p, const AS_GLOBAL double* rt, int KS, int kk,
                                            double* t, double* prem) {
    const double layerratio = rt[RT_LAYER * KS + kk];
    double tn[NSED];
#pragma unroll
    for (int s = 0; s < NSED; s++) tn[s] = 0.0;
    double CaCO3_ratio = 0.0;
    if (p.caco3) {
        const double f1 = rt[RT_CAL1 * KS + kk];
        const double f2 = rt[RT_CAL2 * KS + kk];
        CaCO3_ratio = 1.0 - ((1.0 - t[S_CACO3FR2]) * f1 + t[S_CACO3FR2] * f2);
        tn[S_CACO3FR2] = t[S_CACO3FR2] > NULLSMALL ? (1.0 - f2) * t[S_CACO3FR2] / CaCO3_ratio : 0.0;
    }
    const double f1 = rt[RT_POC1 * KS + kk];
    const double f2 = rt[RT_POC2 * KS + kk];
    const double POC_ratio = 1.0 - ((1.0 - t[S_POCFR2]) * f1 + t[S_POCFR2] * f2);
    tn[S_POCFR2] = t[S_POCFR2] > NULLSMALL ? (1.0 - f2) * t[S_POCFR2] / POC_ratio : 0.0;
    tn[S_POC] = t[S_POC] * layerratio * POC_ratio;
    tn[S_POP] = t[S_POP] * layerratio * POC_ratio;
    tn[S_CACO3] = p.caco3 ? t[S_CACO3] * layerratio * CaCO3_ratio : 0.0;
#pragma unroll
    for (int s = 0; s < NSED; s++) {
        prem[s] = (layerratio * t[s] - tn[s]);
        t[s] = tn[s];
    }
}

// per-level inputs of the column march
struct Lvl {
    double O2, domc, domp, p[NSED];
};

// Particle streams of the march.  Source level k (processed in the
// reference's order, k descending) starts a stream holding the flux that
// leaves it; the stream is stepped through the levels below it until it
// settles at min_k (NEWP) or reaches the sea floor.  At target level kk the
// live streams are the sources kk+1 .. kk+W-1 (the host checks that no
// stream spans W levels or more), source kk+1+j in slot j; after each level
// the streams move down one slot, so every slot index is a compile-time
// constant, the streams stay in registers and the level loop is not unrolled
// (an unrolled march, W copies of W-1 stream steps, was ~120 KB of code:
// more than the instruction cache two CUs share, and it evicted the flux
// kernel running beside it).  The contributions to a level therefore arrive
// in the reference's order: DOM first, then sources from the top down, then
// the sediment return, then uptake.
template <int W>
struct Streams {
    double t[W][NSED];
    int mk[W];  // settling level of the stream, -2 for a free slot
};

template <typename R, int W>
__global__ __launch_bounds__(64) void k_bgc_col(const MemberDev* __restrict__ mems, const int32_t* __restrict__ list) {
    const BlockXYZ B = xcd_blk();
    const MemberDev& M = mems[list[B.y]];
    const GridDev& g = M.g;
    const int w = B.x * 64 + threadIdx.x;
    if (w >= g.nwet) return;
    const BgcParams& p = M.bp;
    const int col = gp(g.wetsorted)[w];
    const int maxi = g.maxi, maxk = g.maxk;
    const int kb = k1at(g, col % maxi + 1, col / maxi + 1);
    const size_t plane = (size_t)g.ncol, ncell = (size_t)g.ncell;
    const int KS = maxk + 2;
    const AS_GLOBAL double* __restrict__ Dbot = gp(g.mk) + MK_DBOT * KS;
    const AS_GLOBAL double* __restrict__ Dmid = gp(g.mk) + MK_DMID * KS;
    const AS_GLOBAL double* __restrict__ rt = gp(M.rtab);
    const AS_GLOBAL double* __restrict__ ocn = gp(M.ocn);  // plane 0 = T, 1 = S, then species planes
    const AS_GLOBAL double* __restrict__ Mm = gp(M.M);
    auto cell = [&](int k) -> size_t { return (size_t)(k - 1) * plane + col; };
    const bool hasO2 = p.slot[B_O2] >= 0, hasDOMC = p.slot[B_DOMC] >= 0, hasDOMP = p.slot[B_DOMP] >= 0;
    const bool hasPO4 = p.slot[B_PO4] >= 0;

    Conv cv;
    make_conv(p, cv);

    // (1) surface carbonate chemistry (biogem.f90:1026-1084)
    double ohm = 0.0;
    if (p.slot[B_DIC] >= 0 && p.slot[B_ALK] >= 0) {
        AS_GLOBAL double* carb = gp(M.carb) + col;
        const size_t cs = cell(maxk);
        const double Tsur = ocn[cs], Ssur = ocn[ncell + cs];
        double cc[NCC];
        calc_carbconst(Dmid[maxk], Tsur, Ssur, cc);
        const double Ca = salt_Ca(Ssur), B = salt_B(Ssur), SO4 = salt_SO4(Ssur), F = salt_F(Ssur);
        const double DIC = ocn[(size_t)p.ocn[B_DIC] * ncell + cs], ALK = ocn[(size_t)p.ocn[B_ALK] * ncell + cs];
        const double PO4 = hasPO4 ? ocn[(size_t)p.ocn[B_PO4] * ncell + cs] : 0.0;
        // FP64 members: the reference's plain iteration; FP32-form members:
        // the same fixed point, accelerated
        constexpr bool fast = sizeof(R) == 4;
        const int it = fast ? calc_carb_fast(DIC, ALK, Ca, PO4, B, SO4, F, cc, carb, plane, p.carb_tol, p.carb_maxit)
                            : calc_carb(DIC, ALK, Ca, PO4, B, SO4, F, cc, carb, plane, p.carb_tol, p.carb_maxit);
        if (it < 0)
            atomicAdd(&M.diag[1], 1);
        else
            atomicAdd(&M.diag[2], it);
        if (fast)
            calc_carb_RF0_fast(DIC, ALK, PO4, B, SO4, F, cc, carb, plane, p.carb_tol, p.carb_maxit);
        else
            calc_carb_RF0(DIC, ALK, PO4, B, SO4, F, cc, carb, plane, p.carb_tol, p.carb_maxit);
        ohm = carb[IC_OHM_CAL * plane];
    }

    // air-sea exchange (biogem.f90:1044-1046, 1605-1614) with the density of
    // biogem_climate (fun_calc_rho, gem_util.f90:670-688; biogem.f90:2187)
    double fair[2] = {0.0, 0.0};
    if (p.airsea) {
        const size_t cs = cell(maxk);
        const double T = ocn[cs], S = ocn[ncell + cs];
        const double TC = T - ZEROC;
        const double rho = 1000.0 + (0.7968 * S - 0.0559 * TC - 0.0063 * (TC * TC) + 3.7315E-05 * (TC * (TC * TC)));
        const int J = col / maxi + 1;
        const double A = cell_area(gp(g.mj) + MJ_SV * (g.maxj + 2), maxi, J);
        const double dD = gp(g.mk)[MK_DD * KS + maxk];
        const double u = gp(M.wspeed)[col];
        const AS_GLOBAL double* atm = gp(M.atm);
        if (p.slot[B_DIC] >= 0 && p.slot[B_ALK] >= 0) {
            const AS_GLOBAL double* carb = gp(M.carb) + col;
            const double sol = solconst(-60.2409, 93.4517, 23.3585, 0.023517, -0.023656, 0.0047036, false, T, S, rho);
            const double pv = piston(2073.1, 125.62, 3.6276, 0.043219, T, u, p.gas_a);
            const double CO2 = carb[IC_CONC_CO2 * plane], RF0 = carb[IC_RF0 * plane];
            double atmv = sol * atm[col], ocnv = CO2, buff;
            if (RF0 > NULLSMALL) {
                buff = 1.0 / (RF0 * CO2 / ocn[(size_t)p.ocn[B_DIC] * ncell + cs]);
            } else {
                ocnv = 0.0;
                atmv = 0.0;
                buff = 1.0;
            }
            fair[0] = ocnatm_flux(pv, A, rho, dD, ocnv, atmv, buff, p.dtyr);
        }
        if (hasO2) {
            const double sol = solconst(-58.3877, 85.8079, 23.8439, -0.034892, 0.015568, -0.0019387, true, T, S, rho);
            const double pv = piston(1953.4, 128.00, 3.9918, 0.050091, T, u, p.gas_a);
            fair[1] = ocnatm_flux(pv, A, rho, dD, ocn[(size_t)p.ocn[B_O2] * ncell + cs], sol * atm[plane + col], 1.0,
                                  p.dtyr);
        }
        gp(M.fatm)[col] = fair[0];
        gp(M.fatm)[plane + col] = fair[1];
    }

    // (2) PO4 uptake of this column (biogem_box.f90:346-1515): the same
    //     particulate and dissolved change at every level from kstart up
    int kstart = maxk + 1;
    double upart[NSED] = {0.0, 0.0, 0.0, 0.0, 0.0}, upt[4] = {0.0, 0.0, 0.0, 0.0}, domC = 0.0, domP = 0.0;
    if (hasPO4) {
        int k_mld = kb;
        const double mld = gp(M.mld)[col];
        for (int k = maxk; k >= 1; k--)
            if (Dbot[k] >= mld) {
                k_mld = k;
                break;
            }
        kstart = k_mld < kb ? kb : k_mld;
        const double lPO4 = ocn[(size_t)p.ocn[B_PO4] * ncell + cell(maxk)];
        const double kPO4 = lPO4 / (lPO4 + p.c0_PO4);
        const double ficefree = (1.0 - 0.0);
        const double kI = gp(M.light)[col];
        double dPO4 = 0.0;
        if (lPO4 > NULLSMALL) dPO4 = p.dtyr * ficefree * kI * kPO4 * p.k0_PO4;
        double DOMfrac = p.DOMfrac;
        double DOMtotal = DOMfrac + 0.0;
        if (DOMtotal > 1.0) {
            DOMfrac = DOMfrac / DOMtotal;
            DOMtotal = 1.0;
        }
        double rCaCO3 = 0.0;
        if (p.caco3 && ohm > 1.0) rCaCO3 = (1.0 - DOMtotal) * p.red_POC_CaCO3 * pow(ohm - 1.0, p.red_POC_CaCO3_pP);
        const double red_POC_POP = 1.0 / p.red_POP_POC;
        const double kPO4f = lPO4 / (lPO4 + p.POC_c0frac2);
        const double fr2 = (1.0 - kPO4f) * p.POC_dfrac2 + p.POC_frac2;
        double bp[NSED];
        bp[S_POC] = p.red_POP_POC * dPO4;
        bp[S_CACO3] = p.caco3 ? rCaCO3 * bp[S_POC] : 0.0;
        bp[S_POP] = red_POC_POP * bp[S_POC];
        for (int s = 0; s < 3; s++) {
            if (s == S_CACO3 && !p.caco3) continue;
            for (int q = 0; q < 4; q++)
                if (cv.c[s][q] != 0.0) upt[q] = upt[q] + cv.c[s][q] * bp[s];
        }
        domC = 1.0 * DOMfrac * bp[S_POC];
        domP = 1.0 * DOMfrac * bp[S_POP];
        bp[S_POC] = bp[S_POC] - (domC + 0.0);
        bp[S_POP] = bp[S_POP] - (domP + 0.0);
        bp[S_CACO3] = bp[S_CACO3] - (0.0 + 0.0);
        bp[S_POCFR2] = fr2;
        bp[S_CACO3FR2] = p.caco3 ? p.CaCO3_frac2 : 0.0;
        for (int s = 0; s < NSED; s++) upart[s] = bp[s];
    }

    // (3) the march, surface to k1
    AS_GLOBAL double* __restrict__ part = gp(M.part);
    const double pscale = gp(M.scal)[SC_PSCALE];  // the last coupling's ratio, pending on part
    AS_GLOBAL double* __restrict__ vd = gp(M.vd);
    const AS_GLOBAL double* __restrict__ pO2 = hasO2 ? ocn + (size_t)p.ocn[B_O2] * ncell : nullptr;
    const AS_GLOBAL double* __restrict__ pDC = hasDOMC ? ocn + (size_t)p.ocn[B_DOMC] * ncell : nullptr;
    const AS_GLOBAL double* __restrict__ pDP = hasDOMP ? ocn + (size_t)p.ocn[B_DOMP] * ncell : nullptr;
    const int klim = (p.dtyr * p.sinkrate <= p.dsc) ? kb : maxk;
    const double ratio = (p.DOMlifetime > p.dtyr) ? p.dtyr / p.DOMlifetime : 1.0;
    const bool pfirst = hasDOMP && p.slot[B_DOMP] < p.slot[B_DOMC];
    const double Mkb = Mm[cell(kb)], rMkb = gp(M.rM)[cell(kb)];
    double sprev[NSED];
#pragma unroll
    for (int s = 0; s < NSED; s++) sprev[s] = gp(M.settle)[(size_t)s * plane + col];
    double settle[NSED] = {0.0, 0.0, 0.0, 0.0, 0.0};
    auto load = [&](int k, Lvl& v) {
        const size_t c = cell(k);
        v.O2 = hasO2 ? pO2[c] : 0.0;
        v.domc = hasDOMC ? pDC[c] : 0.0;
        v.domp = (hasDOMC && hasDOMP) ? pDP[c] : 0.0;
#pragma unroll
        for (int s = 0; s < NSED; s++) v.p[s] = k >= klim ? pscale * (part[(size_t)s * ncell + c] + 0.0) : 0.0;
    };
    // settled flux at k1 (the kk = k1 term of the settle loop)
    auto settle_add = [&](const double* t) {
#pragma unroll
        for (int s = 0; s < NSED; s++)
            settle[s] = settle[s] + ((s == S_POCFR2 || s == S_CACO3FR2) ? t[s] : Mkb * t[s]);
    };
    // one target level; `streams` steps the live streams in source order
    auto level = [&](int kk, const Lvl& cur, auto&& streams, auto&& push) {
        const double rf = redfield(p, cur.O2);
        double v[4] = {0.0, 0.0, 0.0, 0.0};
        double vdc = 0.0, vdp = 0.0;
        if (hasDOMC) {
            double bp[NSED] = {0.0, 0.0, 0.0, 0.0, 0.0};
            if (cur.domc > NULLSMALL) {
                for (int q = 0; q < 2; q++) {
                    const bool isC = (q == 0) != pfirst;
                    if (!isC && !hasDOMP) continue;
                    const double x = isC ? cur.domc : cur.domp;
                    if (isC) {
                        bp[S_POC] = bp[S_POC] + 1.0 * ratio * x;
                        vdc = vdc - ratio * x;
                    } else {
                        bp[S_POP] = bp[S_POP] + 1.0 * ratio * x;
                        vdp = vdp - ratio * x;
                    }
                }
            }
            add_remin(p, cv, rf, bp, v);
        }
        double np[NSED] = {0.0, 0.0, 0.0, 0.0, 0.0};
        streams(kk, rf, v, np);
        // source level kk itself
        if (kk >= klim) {
            double tot = 0.0;
            tot = tot + cur.p[S_POC];
            if (p.caco3) tot = tot + cur.p[S_CACO3];
            if (tot > NULLSMALL) {
                if (kk == kb)
                    settle_add(cur.p);
                else
                    push(cur.p, gp(M.mink)[kk] >= kb ? gp(M.mink)[kk] : kb - 1);
            }
        }
        // sediment return of the previous step's rain (biogem.f90:873-923, 937-942)
        if (kk == kb) {
            double fsed[4] = {0.0, 0.0, 0.0, 0.0};
#pragma unroll
            for (int s = 0; s < NSED; s++) {
                if (s == S_CACO3 && !p.caco3) continue;
#pragma unroll
                for (int q = 0; q < 4; q++)
                    if (cv.c[s][q] != 0.0) fsed[q] = fsed[q] + (rf * cv.c[s][q]) * sprev[s];
            }
#pragma unroll
            for (int q = 0; q < 4; q++) v[q] = rMkb * fsed[q] + v[q];
        }
        const size_t c = cell(kk);
        const bool up = hasPO4 && kk >= kstart;
        if (up) {
            vdc = vdc + domC;
            vdp = vdp + domP;
#pragma unroll
            for (int q = 0; q < 4; q++) v[q] = v[q] - upt[q];
        }
        // vdocn = bio_remin + dtyr rM locijk_focn (biogem.f90:1821-1833): the
        // surface takes minus the ocean -> atmosphere flux
        if (p.airsea && kk == maxk) {
            const double dtrM = p.dtyr * gp(M.rM)[c];
            v[B_DIC] = v[B_DIC] + dtrM * (0.0 - fair[0]);
            v[B_O2] = v[B_O2] + dtrM * (0.0 - fair[1]);
        }
#pragma unroll
        for (int s = 0; s < NSED; s++) part[(size_t)s * ncell + c] = up ? upart[s] : np[s];
#pragma unroll
        for (int q = 0; q < 4; q++)
            if (p.slot[q] >= 0) vd[(size_t)q * ncell + c] = v[q];
        if (hasDOMC) vd[(size_t)B_DOMC * ncell + c] = vdc;
        if (hasDOMP) vd[(size_t)B_DOMP * ncell + c] = vdp;
    };
    // stepping one live stream through level kk
    auto advance = [&](double* t, int& mk, int kk, double rf, double* v, double* np) {
        double prem[NSED];
        stream_step(p, rt, KS, kk, t, prem);
        add_remin(p, cv, rf, prem, v);
        if (kk == kb && kk >= mk + 1) settle_add(t);
        if (kk == mk) {
#pragma unroll
            for (int s = 0; s < NSED; s++) np[s] = np[s] + t[s];
            mk = -2;
        } else if (kk == kb) {
            mk = -2;
        }
    };

    Lvl cur, nxt;
    load(maxk, cur);
    if (klim == maxk) {
        // particles sink out of the column within one step: one source
        double t1[NSED];
        int mk1 = -2;
        for (int kk = maxk; kk >= kb; kk--) {
            if (kk > kb) load(kk - 1, nxt);
            level(
                kk, cur,
                [&](int k_, double rf, double* v, double* np) {
                    if (mk1 > -2) advance(t1, mk1, k_, rf, v, np);
                },
                [&](const double* t, int mk) {
#pragma unroll
                    for (int s = 0; s < NSED; s++) t1[s] = t[s];
                    mk1 = mk;
                });
            cur = nxt;
        }
    } else {
        // at level kk slot j holds the stream of source kk + 1 + j; after the
        // level every stream moves down one slot and source kk enters slot 0
        // (the stream of source kk + W - 1 has settled by then)
        Streams<W - 1> S;
#pragma unroll
        for (int j = 0; j < W - 1; j++) S.mk[j] = -2;
        for (int kk = maxk; kk >= kb; kk--) {
            if (kk > kb) load(kk - 1, nxt);
            double pt[NSED] = {0.0, 0.0, 0.0, 0.0, 0.0};
            int pmk = -2;
            level(
                kk, cur,
                [&](int k_, double rf, double* v, double* np) {
#pragma unroll
                    for (int j = W - 2; j >= 0; j--)  // sources from the top down
                        if (S.mk[j] > -2) advance(S.t[j], S.mk[j], k_, rf, v, np);
                },
                [&](const double* t, int mk) {
#pragma unroll
                    for (int s = 0; s < NSED; s++) pt[s] = t[s];
                    pmk = mk;
                });
#pragma unroll
            for (int j = W - 2; j >= 1; j--) {
#pragma unroll
                for (int s = 0; s < NSED; s++) S.t[j][s] = S.t[j - 1][s];
                S.mk[j] = S.mk[j - 1];
            }
#pragma unroll
            for (int s = 0; s < NSED; s++) S.t[0][s] = pt[s];
            S.mk[0] = pmk;
            cur = nxt;
        }
    }
#pragma unroll
    for (int s = 0; s < NSED; s++) gp(M.settle)[(size_t)s * plane + col] = settle[s];
}

// Column kernels of the coupling: one block per (member, 64 wet columns).
// Thread (c, g) of the block owns column c's cells at levels g+1, g+5, ...
// (CPL_LG level groups, NJ cells), so each cell's data is loaded once and the
// tracers are taken in groups of `tg`: every thread forms its cells'
// products for the group, the products go to LDS, and one thread per
// (column, tracer) adds its column from k1 up, in the reference's order.
constexpr int CPL_THREADS = 256;
constexpr int CPL_COLS = 64;
constexpr int CPL_LG = CPL_THREADS / CPL_COLS;         // level groups
constexpr int CPL_MAXTG = 1;                            // tracers per LDS round (1: fewer registers, more waves)
constexpr size_t CPL_LDS_CAP = 64 * 1024;

__host__ __device__ inline int cpl_tg(int maxk) {
    int tg = (int)(CPL_LDS_CAP / ((size_t)maxk * CPL_COLS * 8));
    return tg > CPL_MAXTG ? CPL_MAXTG : (tg < 1 ? 1 : tg);
}

struct CplTile {
    int col[CPL_COLS], kb[CPL_COLS], slot[CPL_COLS];
};

__device__ __forceinline__ void cpl_tile(const GridDev& g, int w0, CplTile& T) {
    const int t = threadIdx.x;
    if (t < CPL_COLS) {
        const int w = w0 + t;
        const int col = w < g.nwet ? gp(g.wetsorted)[w] : 0;
        T.col[t] = col;
        T.kb[t] = w < g.nwet ? k1at(g, col % g.maxi + 1, col / g.maxi + 1) : g.maxk + 1;
        T.slot[t] = w < g.nwet ? gp(g.vpos)[w] : 0;
    }
    __syncthreads();
}

// column partial sums of the state after transport: sum 0 is S
// (biogem.f90:2016-2025), sum q >= 1 is tracer q + 1 (1984-1997)
template <typename R, int NJ, int TG>  // NJ cells per thread (NJ * CPL_LG >= maxk), TG >= tg
__global__ __launch_bounds__(CPL_THREADS) void k_couple_sums(const MemberDev* __restrict__ mems,
                                                             const int32_t* __restrict__ list) {
    extern __shared__ double prod[];  // [tg][maxk][CPL_COLS]
    __shared__ CplTile T;
    const BlockXYZ B = xcd_blk();
    const MemberDev& M = mems[list[B.y]];
    const GridDev& g = M.g;
    const int w0 = B.x * CPL_COLS;
    if (w0 >= g.nwet) return;        // uniform over the block
    if (*gp(M.sfresh)) return;        // the transport step already formed them
    cpl_tile(g, w0, T);
    const int c = threadIdx.x % CPL_COLS, lg = threadIdx.x / CPL_COLS;
    const int maxk = g.maxk, L = M.L, nsum = L - 1, tg = cpl_tg(maxk) < TG ? cpl_tg(maxk) : TG;
    const int col = T.col[c], kb = T.kb[c];
    const size_t plane = (size_t)g.ncol, ncell = (size_t)g.ncell, tstride = (size_t)maxk * plane;
    const int par = *gp(M.parity);  // the transport step has flipped: st[par] is the new state
    const void* st = M.st[par];
    const int16_t* lo = M.lo[par];
    const double rSold = 1.0 / gp(M.scal)[SC_SMEAN];
    double os[NJ], mm[NJ];
#pragma unroll
    for (int j = 0; j < NJ; j++) {
        const int k = 1 + lg + CPL_LG * j;
        if (k <= maxk && k >= kb) {
            const size_t cc = (size_t)(k - 1) * plane + col;
            os[j] = gp(M.ocn)[ncell + cc];
            mm[j] = gp(M.M)[cc];
        }
    }
    for (int q0 = 0; q0 < nsum; q0 += tg) {
        // all loads of the group first, then the products
        typename StateIO<R>::Raw x[TG][NJ];
        double a[NJ];
#pragma unroll
        for (int qt = 0; qt < TG; qt++) {
            const int q = q0 + qt, l = q == 0 ? 1 : q + 1;
            if (qt < tg && q < nsum) {
#pragma unroll
                for (int j = 0; j < NJ; j++) {
                    const int k = 1 + lg + CPL_LG * j;
                    if (k <= maxk && k >= kb) {
                        const size_t cc = (size_t)(k - 1) * plane + col;
                        x[qt][j] = StateIO<R>::ldr(st, lo, (size_t)l * tstride + cc);
                        if (q == 0) a[j] = gp(g.V)[cc];
                    }
                }
            }
        }
#pragma unroll
        for (int qt = 0; qt < TG; qt++) {
            const int q = q0 + qt;
            if (qt < tg && q < nsum) {
#pragma unroll
                for (int j = 0; j < NJ; j++) {
                    const int k = 1 + lg + CPL_LG * j;
                    if (k <= maxk && k >= kb) {
                        const double xv = StateIO<R>::val(x[qt][j], lo != nullptr);
                        double v;
                        if (q == 0) {
                            const double Sn = xv + SALN0 + 0.0;
                            v = Sn * a[j];
                        } else {
                            const double lv = xv * os[j] * rSold;
                            v = lv * mm[j];
                        }
                        prod[((size_t)qt * maxk + (k - 1)) * CPL_COLS + c] = v;
                    }
                }
            }
        }
        __syncthreads();
        {
            const int qt = lg;  // one thread per (column, sum of the group)
            const int q = q0 + qt;
            if (qt < tg && q < nsum && w0 + c < g.nwet) {
                double acc = 0.0;
                for (int k = kb; k <= maxk; k++) acc += prod[((size_t)qt * maxk + (k - 1)) * CPL_COLS + c];
                gp(M.colsum)[(size_t)(q == 0 ? 0 : q + 1) * g.nwet + T.slot[c]] = acc;
            }
        }
        __syncthreads();
    }
}

// global sums in the reference's column order (partialtot(n) summed over
// n = 1..n_vocn, biogem.f90:1927-2031), then the coupling scalars.  The
// block stages a chunk of columns of every partial-sum row in LDS, then each
// sum is one thread's sequential pass over its row of the chunk; chunks go in
// column order, so every sum is added in the reference's order.
constexpr int RED_THREADS = 1024;  // one block per member: many loads in flight per chunk
constexpr int RED_LDS_BYTES = 128 * 1024;
constexpr int RED_MAXL = 256;

__global__ __launch_bounds__(RED_THREADS) void k_couple_reduce(const MemberDev* __restrict__ mems,
                                                               const int32_t* __restrict__ list, int lds_bytes) {
    extern __shared__ double rows[];
    __shared__ double s_val[2 * RED_MAXL];
    const MemberDev& M = mems[list[blockIdx.x]];
    const int nwet = M.g.nwet, L = M.L;
    // sum q: 0 -> S (scaled by 1/Vtot), 1 + (l-2) -> NEW_l, L-1 + (l-2) -> OLD_l
    const int nsum = 1 + 2 * (L - 2);
    auto src = [&](int q) -> const AS_GLOBAL double* {
        if (q == 0) return gp(M.colsum);
        if (q < L - 1) return gp(M.colsum) + (size_t)(q + 1) * nwet;
        return gp(M.colold) + (size_t)(q - L + 3) * nwet;
    };
    int ch = lds_bytes / (8 * nsum);   // columns per chunk, odd (LDS banks)
    if (ch > nwet) ch = nwet;
    if ((ch & 1) == 0) ch -= 1;
    if (ch < 1) ch = 1;
    const int t = threadIdx.x;
    const double sc = M.g.rVtot;
    double tot0 = 0.0, tot1 = 0.0;  // sums t and t + RED_THREADS
    for (int c0 = 0; c0 < nwet; c0 += ch) {
        const int n = nwet - c0 < ch ? nwet - c0 : ch;
        // stage nsum x n values, eight loads in flight per thread
        const int tot = nsum * n;
        for (int b0 = t; b0 < tot; b0 += RED_THREADS * 8) {
            double v[8];
            int at[8];
#pragma unroll
            for (int u = 0; u < 8; u++) {
                const int idx = b0 + u * RED_THREADS;
                if (idx < tot) {
                    const int q = idx / n, i = idx - q * n;
                    v[u] = src(q)[c0 + i];
                    at[u] = q * ch + i;
                }
            }
#pragma unroll
            for (int u = 0; u < 8; u++)
                if (b0 + u * RED_THREADS < tot) rows[at[u]] = v[u];
        }
        __syncthreads();
        // each sum adds its columns in order; the LDS reads of eight columns
        // are issued before the first of their adds
        auto run = [&](const double* r, double tot, double f, bool scaled) {
            int i = 0;
            for (; i + 8 <= n; i += 8) {
                double v[8];
#pragma unroll
                for (int u = 0; u < 8; u++) v[u] = r[i + u];
#pragma unroll
                for (int u = 0; u < 8; u++) tot += scaled ? v[u] * f : v[u];
            }
            for (; i < n; i++) tot += scaled ? r[i] * f : r[i];
            return tot;
        };
        if (t < nsum) tot0 = run(rows + t * ch, tot0, sc, t == 0);
        if (t + RED_THREADS < nsum) tot1 = run(rows + (t + RED_THREADS) * ch, tot1, 1.0, false);
        __syncthreads();
    }
    if (t < nsum) s_val[t] = tot0;
    if (t + RED_THREADS < nsum) s_val[t + RED_THREADS] = tot1;
    __syncthreads();
    for (int l = 2 + t; l < L; l += RED_THREADS) {
        const double nw = s_val[l - 1];
        const double rnew = fabs(nw) < NULLSMALL ? 0.0 : 1.0 / nw;
        gp(M.scal)[SC_FACT0 + l] = s_val[L - 3 + l] * rnew;
    }
    if (t == 0) {
        const double Sold = gp(M.scal)[SC_SMEAN];
        const double Snew = s_val[0];
        const double Sratio = Snew / Sold;
        gp(M.scal)[SC_SOLD] = Sold;
        gp(M.scal)[SC_SNEW] = Snew;
        gp(M.scal)[SC_SMEAN] = Snew;
        gp(M.scal)[SC_SRATIO] = Sratio;
        gp(M.scal)[SC_RSRATIO] = 1.0 / Sratio;
    }
}

// biogem_tracercoupling's rescale (biogem.f90:2001-2061): tracers l >= 3
// (2033-2046) with the OLD partial sums of the next coupling, then T and S
// (2001-2013), the new cell masses (2049-2052) and the particulates
// (2042-2043).  The block owns its cells, so the old salinity and masses the
// tracers need are read before the same threads rewrite them.
template <typename R, int NJ, int TG>  // NJ cells per thread (NJ * CPL_LG >= maxk), TG >= tg
__global__ __launch_bounds__(CPL_THREADS) void k_couple_apply(const MemberDev* __restrict__ mems,
                                                              const int32_t* __restrict__ list) {
    extern __shared__ double prod[];  // [tg][maxk][CPL_COLS]
    __shared__ CplTile T;
    const BlockXYZ B = xcd_blk();
    const MemberDev& M = mems[list[B.y]];
    const GridDev& g = M.g;
    const int w0 = B.x * CPL_COLS;
    if (w0 >= g.nwet) return;  // uniform over the block
    cpl_tile(g, w0, T);
    const BgcParams& p = M.bp;
    const int c = threadIdx.x % CPL_COLS, lg = threadIdx.x / CPL_COLS;
    const int maxk = g.maxk, L = M.L, tg = cpl_tg(maxk) < TG ? cpl_tg(maxk) : TG;
    const int col = T.col[c], kb = T.kb[c];
    const size_t plane = (size_t)g.ncol, ncell = (size_t)g.ncell, tstride = (size_t)maxk * plane;
    const int par = *gp(M.parity);
    void* st = M.st[par];
    int16_t* lo = M.lo[par];
    AS_GLOBAL double* __restrict__ ocn = gp(M.ocn);
    const double rSold = 1.0 / gp(M.scal)[SC_SOLD];
    const double Snew = gp(M.scal)[SC_SNEW], Sratio = gp(M.scal)[SC_SRATIO], rSratio = gp(M.scal)[SC_RSRATIO];
    // per cell: old salinity, Snew / new salinity, new mass
    double so[NJ], rs[NJ], mn[NJ];
    {
        typename StateIO<R>::Raw sr[NJ];
#pragma unroll
        for (int j = 0; j < NJ; j++) {
            const int k = 1 + lg + CPL_LG * j;
            if (k <= maxk && k >= kb) {
                const size_t cc = (size_t)(k - 1) * plane + col;
                so[j] = ocn[ncell + cc];
                sr[j] = StateIO<R>::ldr(st, lo, tstride + cc);
                mn[j] = gp(M.M)[cc];
            }
        }
#pragma unroll
        for (int j = 0; j < NJ; j++) {
            const int k = 1 + lg + CPL_LG * j;
            if (k <= maxk && k >= kb) {
                const double snew = StateIO<R>::val(sr[j], lo != nullptr) + SALN0 + 0.0;
                rs[j] = Snew / snew;
                mn[j] = rSratio * mn[j];
            }
        }
    }
    for (int l0 = 2; l0 < L; l0 += tg) {
        // all loads of the group first, then the rescale
        typename StateIO<R>::Raw lv[TG][NJ];
        double vv[TG][NJ];
        int bs[TG];
#pragma unroll
        for (int lt = 0; lt < TG; lt++) {
            const int l = l0 + lt;
            int b = -1;
            for (int q = 0; q < NBGC; q++)
                if (p.slot[q] == l) b = q;
            bs[lt] = b;
            const bool hasv = M.bgc && b >= 0;
            if (lt < tg && l < L) {
#pragma unroll
                for (int j = 0; j < NJ; j++) {
                    const int k = 1 + lg + CPL_LG * j;
                    if (k <= maxk && k >= kb) {
                        const size_t cc = (size_t)(k - 1) * plane + col;
                        lv[lt][j] = StateIO<R>::ldr(st, lo, (size_t)l * tstride + cc);
                        vv[lt][j] = hasv ? gp(M.vd)[(size_t)b * ncell + cc] : 0.0;
                    }
                }
            }
        }
#pragma unroll
        for (int lt = 0; lt < TG; lt++) {
            const int l = l0 + lt;
            if (lt < tg && l < L) {
                const int b = bs[lt];
                const double fact = gp(M.scal)[SC_FACT0 + l];
#pragma unroll
                for (int j = 0; j < NJ; j++) {
                    const int k = 1 + lg + CPL_LG * j;
                    if (k <= maxk && k >= kb) {
                        const size_t cc = (size_t)(k - 1) * plane + col;
                        const double x = StateIO<R>::val(lv[lt][j], lo != nullptr) * so[j] * rSold;
                        double v = fact * x + vv[lt][j];
                        v = Sratio * v;
                        if (b >= 0) ocn[(size_t)p.ocn[b] * ncell + cc] = v;
                        StateIO<R>::st(st, lo, (size_t)l * tstride + cc, rs[j] * v);
                        prod[((size_t)lt * maxk + (k - 1)) * CPL_COLS + c] = v * mn[j];
                    }
                }
            }
        }
        __syncthreads();
        {
            const int lt = lg;
            const int l = l0 + lt;
            if (lt < tg && l < L && w0 + c < g.nwet) {
                double acc = 0.0;
                for (int k = kb; k <= maxk; k++) acc += prod[((size_t)lt * maxk + (k - 1)) * CPL_COLS + c];
                gp(M.colold)[(size_t)l * g.nwet + T.slot[c]] = acc;
            }
        }
        __syncthreads();
    }
    // the state changes below: the next coupling's sums come from the next transport step
    if (B.x == 0 && threadIdx.x == 0) {
        *gp(M.sfresh) = 0;
        // the particulates' rescale (biogem.f90:2042-2043) is left to their next reader
        gp(M.scal)[SC_PSCALE] = Sratio;
    }
    // T, S, masses and particulates of the thread's cells
#pragma unroll
    for (int j = 0; j < NJ; j++) {
        const int k = 1 + lg + CPL_LG * j;
        if (k <= maxk && k >= kb) {
            const size_t cc = (size_t)(k - 1) * plane + col;
            const auto rT = StateIO<R>::ldr(st, lo, cc);
            const auto rS = StateIO<R>::ldr(st, lo, tstride + cc);
            const double lvT = StateIO<R>::val(rT, lo != nullptr) + ZEROC + 0.0;
            const double lvS = StateIO<R>::val(rS, lo != nullptr) + SALN0 + 0.0;
            ocn[cc] = lvT;
            ocn[ncell + cc] = lvS;
            StateIO<R>::st(st, lo, cc, lvT - ZEROC);
            StateIO<R>::st(st, lo, tstride + cc, lvS - SALN0);
            gp(M.M)[cc] = mn[j];
            // 1/M is read only at k1 (the sediment return) and at the surface (the air-sea flux)
            if (k == kb || k == maxk) gp(M.rM)[cc] = Sratio * gp(M.rM)[cc];
        }
    }
}

// The 17 carbonate constants (sub_calc_carbconst, gem_carbchem.f90:59-268)
// at the surface cell of every wet column from the current ocn T and S, as
// the column step evaluates them (a download for checking them).
__global__ __launch_bounds__(64) void k_carbconst_surface(const MemberDev* __restrict__ mems, int32_t m,
                                                          double* __restrict__ out) {
    const MemberDev& M = mems[m];
    const GridDev& g = M.g;
    const int w = blockIdx.x * 64 + threadIdx.x;
    if (w >= g.nwet) return;
    const int col = gp(g.wetsorted)[w];
    const size_t plane = (size_t)g.ncol, ncell = (size_t)g.ncell, cs = (size_t)(g.maxk - 1) * plane + col;
    double cc[NCC];
    calc_carbconst(gp(g.mk)[MK_DMID * (g.maxk + 2) + g.maxk], gp(M.ocn)[cs], gp(M.ocn)[ncell + cs], cc);
    for (int q = 0; q < NCC; q++) out[(size_t)q * plane + col] = cc[q];
}

void launch_carbconst_surface(const MemberDev* mems, int32_t m, int nwet, double* out, hipStream_t s) {
    hipLaunchKernelGGL(k_carbconst_surface, dim3((nwet + 63) / 64), dim3(64), 0, s, mems, m, out);
}

// cpl_flux_ocnatm (atchem.f90:306-320) after BIOGEM's interface update
// (biogem.f90:1719-1727), then step_atchem (atchem.f90:63-158) with no
// decay, CH4 or 14C sources: one block per member.  Each cell's new partial
// pressure is staged as its mole content, the contents are summed in the
// reference's SUM order (i fastest) by one lane per gas, and the atmosphere
// is left well mixed at 273.15 K (atchem_data.f90:248).
constexpr int ATM_THREADS = 256;
constexpr int ATM_LDS_COLS = 4096;  // columns staged in LDS; larger grids sum from global memory
__global__ __launch_bounds__(ATM_THREADS) void k_atchem(const MemberDev* __restrict__ mems,
                                                        const int32_t* __restrict__ list, int lds_cols) {
    extern __shared__ double stage[];  // [2][lds_cols]
    __shared__ double tot[3];
    const MemberDev& M = mems[list[blockIdx.x]];
    if (!M.bgc || !M.bp.airsea) return;  // uniform over the block
    const GridDev& g = M.g;
    const int maxi = g.maxi, ncol = g.ncol;
    const AS_GLOBAL double* sv = gp(g.mj) + MJ_SV * (g.maxj + 2);
    AS_GLOBAL double* atm = gp(M.atm);
    const AS_GLOBAL double* fatm = gp(M.fatm);
    const bool lds = ncol <= lds_cols;  // the launch sized the LDS for lds_cols columns
    const double dts = M.bp.dtyr * CONV_YR_S;
    for (int c = threadIdx.x; c < ncol; c += ATM_THREADS) {
        const double A = cell_area(sv, maxi, c / maxi + 1);
        const double V = ATM_TH * A;
        const double cam = V / (CONV_PA_ATM * CONST_R_SI * ZEROC);
        const double cma = 1.0 / cam;
#pragma unroll
        for (int q = 0; q < 2; q++) {
            const double sfx = (1.0 / A) * CONV_S_YR * fatm[(size_t)q * ncol + c];
            const double sum = 0.0 + dts * sfx;
            const double a = atm[(size_t)q * ncol + c] + cma * A * sum + cma * 0.0;
            if (lds)
                stage[q * ncol + c] = cam * a;
            else
                atm[(size_t)q * ncol + c] = cam * a;
        }
    }
    __syncthreads();
    if (threadIdx.x < 3) {
        const int q = threadIdx.x;
        double acc = 0.0;
        if (q == 2) {
            // a cell's volume depends only on its row: formed once per row,
            // added once per cell in the same order (i fastest)
            for (int j = 0; j < ncol / maxi; j++) {
                const double v = ATM_TH * cell_area(sv, maxi, j + 1);
                for (int i = 0; i < maxi; i++) acc += v;
            }
        } else if (lds) {
#pragma unroll 8
            for (int c = 0; c < ncol; c++) acc += stage[q * ncol + c];
        } else {
#pragma unroll 8
            for (int c = 0; c < ncol; c++) acc += atm[(size_t)q * ncol + c];
        }
        tot[q] = acc;
    }
    __syncthreads();
    const double p0 = (tot[0] / tot[2]) * CONV_PA_ATM * CONST_R_SI * ZEROC;
    const double p1 = (tot[1] / tot[2]) * CONV_PA_ATM * CONST_R_SI * ZEROC;
    for (int c = threadIdx.x; c < ncol; c += ATM_THREADS) {
        atm[c] = p0;
        atm[(size_t)ncol + c] = p1;
    }
}

void launch_atchem(const MemberDev* mems, const int32_t* list, int nlist, int ncolmax, hipStream_t s) {
    const int lds_cols = ncolmax <= ATM_LDS_COLS ? ncolmax : 0;
    hipLaunchKernelGGL(k_atchem, dim3(nlist), dim3(ATM_THREADS), (size_t)2 * lds_cols * sizeof(double), s, mems, list,
                       lds_cols);
}

// ------------------------------------------------------------ launchers
template <typename R>
static void launch_col(int W, dim3 grid, const MemberDev* mems, const int32_t* list, hipStream_t s) {
    if (W <= 4)
        hipLaunchKernelGGL((k_bgc_col<R, 4>), grid, dim3(64), 0, s, mems, list);
    else if (W <= 5)
        hipLaunchKernelGGL((k_bgc_col<R, 5>), grid, dim3(64), 0, s, mems, list);
    else if (W <= 6)
        hipLaunchKernelGGL((k_bgc_col<R, 6>), grid, dim3(64), 0, s, mems, list);
    else if (W <= 8)
        hipLaunchKernelGGL((k_bgc_col<R, 8>), grid, dim3(64), 0, s, mems, list);
    else
        hipLaunchKernelGGL((k_bgc_col<R, 16>), grid, dim3(64), 0, s, mems, list);
}

void launch_bio(const MemberDev* mems, const int32_t* list, int nlist, int nwetmax, int W, int prec, hipStream_t s) {
    const dim3 gcol((nwetmax + 63) / 64, nlist);
    if (prec == 0)
        launch_col<double>(W, gcol, mems, list, s);
    else
        launch_col<float>(W, gcol, mems, list, s);
}

template <typename R>
// phase 1: column sums (if stale) and the global sums; phase 2: the rescale;
// phase 0: both; phase 3: the global sums alone (right after a transport step)
static void launch_coupling_t(const MemberDev* mems, const int32_t* list, int nlist, int nwetmax, int maxkmax,
                              size_t lds_cap, int phase, hipStream_t s) {
    const dim3 grid((nwetmax + CPL_COLS - 1) / CPL_COLS, nlist);
    // LDS for the largest grid of the batch; every member's tg fits in it
    const size_t lds = (size_t)cpl_tg(maxkmax) * maxkmax * CPL_COLS * 8;
    size_t need = 0;
    for (int k = 1; k <= maxkmax; k++) need = std::max(need, (size_t)cpl_tg(k) * k * CPL_COLS * 8);
    const size_t b = need > lds ? need : lds;
    if (phase != 2) {
        if (phase == 3) {
            // the transport step just formed the column sums
        } else if (maxkmax <= 4 * CPL_LG) {
            hipLaunchKernelGGL((k_couple_sums<R, 4, 1>), grid, dim3(CPL_THREADS), b, s, mems, list);
        } else if (maxkmax <= 8 * CPL_LG) {
            hipLaunchKernelGGL((k_couple_sums<R, 8, 1>), grid, dim3(CPL_THREADS), b, s, mems, list);
        } else {
            hipLaunchKernelGGL((k_couple_sums<R, 12, 1>), grid, dim3(CPL_THREADS), b, s, mems, list);
        }
        // the staging chunk takes what LDS the device grants a block, up to RED_LDS_BYTES
        const int red_lds = (int)std::min<size_t>(RED_LDS_BYTES, lds_cap - 2 * RED_MAXL * sizeof(double));
        hipLaunchKernelGGL(k_couple_reduce, dim3(nlist), dim3(RED_THREADS), (size_t)red_lds, s, mems, list, red_lds);
    }
    if (phase == 1 || phase == 3) return;
    if (maxkmax <= 4 * CPL_LG) {
        hipLaunchKernelGGL((k_couple_apply<R, 4, 1>), grid, dim3(CPL_THREADS), b, s, mems, list);
    } else if (maxkmax <= 8 * CPL_LG) {
        hipLaunchKernelGGL((k_couple_apply<R, 8, 1>), grid, dim3(CPL_THREADS), b, s, mems, list);
    } else {
        hipLaunchKernelGGL((k_couple_apply<R, 12, 1>), grid, dim3(CPL_THREADS), b, s, mems, list);
    }
}

void launch_coupling(const MemberDev* mems, const int32_t* list, int nlist, int nwetmax, int maxkmax, int maxL,
                     int prec, size_t lds_cap, int phase, hipStream_t s) {
    (void)maxL;
    if (prec == 0)
        launch_coupling_t<double>(mems, list, nlist, nwetmax, maxkmax, lds_cap, phase, s);
    else
        launch_coupling_t<float>(mems, list, nlist, nwetmax, maxkmax, lds_cap, phase, s);
}

}  // namespace cgo
