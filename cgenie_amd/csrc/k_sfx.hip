// k_sfx.hip -- EMBM's surface fluxes (surflux, embm.f90:2548-3738) in the
// configuration without ENTS, and the surface slot ts(1:2, :, :, maxk+1)
// that step_goldstein forms from them (goldstein.f90:149-170).
//
// One workgroup per member.  The first pass computes every cell's fluxes
// (the loop of embm.f90:2955-3637 and the ocean-side totals of 3640-3679);
// precipitation goes to LDS.  After a barrier, the second pass gathers each
// wet column's runoff from its land sources in the order the reference's
// main loop adds them (i outer, j inner), so the sum rounds alike, and writes
// the surface slot.  FP64 throughout, each expression in the reference's
// operand order (built with -ffp-contract=off).
#include "cgocn_dev.h"

namespace cgo {
namespace {

// embm_lib.f90:36-157 (PARAMETERs)
constexpr double RHOAIR = 1.25, RHO0 = 1.0E3, RHOAO = 1.25 / 1.0E3;
constexpr double CPA = 1004.0, HLV = 2.501E6, HLS = 2.501E6 + 3.34E5;
constexpr double CO20 = 278.0E-6, CH40 = 700.0E-9, N2O0 = 275.0E-9;
constexpr double ALPHACH4 = 0.036, ALPHAN2O = 0.12;
constexpr double B00 = 2.43414E2, B10 = -3.47968E1, B20 = 1.02790E1;
constexpr double B01 = 2.60065, B11 = -1.62064, B21 = 6.34856E-1;
constexpr double B02 = 4.40272E-3, B12 = -2.26092E-2, B22 = 1.12265E-2;
constexpr double B03 = -2.05237E-5, B13 = -9.67000E-5, B23 = 5.62925E-5;
constexpr double EMO = 0.94 * 5.67E-8, EMA = 0.85 * 5.67E-8;
constexpr double TFREEZ = 0.0;
constexpr double CONST1 = 3.80E-3, CONST2 = 21.87, CONST3 = 265.5, CONST4 = 17.67, CONST5 = 243.5;
constexpr double CONSIC = 2.166;
constexpr double RHMIN = 1.0 / 0.01, RHOOI = 1.0E3 / 913.0, RRHOLF = 1.0 / (913.0 * 3.34E5);
constexpr double M2MM = 1000.0, MM2M = 1.0 / 1000.0;
constexpr double HATMBL2 = 1800.;  // hatmbl(2), embm.f90:1081
// goldstein_lib.f90:49-63 and goldstein.f90:904: the slot's scalings
constexpr double G_RSC = 6.37E6, G_DSC = 5.0E3, G_USC = 0.05, G_RH0SC = 1.0E3, G_CPSC = 3981.1;
constexpr double RFLUXSC = G_RSC / (G_DSC * G_USC * G_RH0SC * G_CPSC);
constexpr double RPMESCO = G_RSC * SALN0 / (G_DSC * G_USC);
constexpr double TOL = 1.0E-10;  // embm.f90:2673
constexpr int ITICE = 21;        // 2637
constexpr int SFX_NT = 256;

__device__ __forceinline__ double fmx(double a, double b) { return b > a ? b : a; }  // MAX(a, b)
__device__ __forceinline__ double fmn(double a, double b) { return b < a ? b : a; }  // MIN(a, b)
__device__ __forceinline__ double p4(double x) {
    const double x2 = x * x;
    return x2 * x2;
}
__device__ __forceinline__ double p3(double x) { return x * (x * x); }

// ch4_func, embm.f90:3771-3777
__device__ double ch4_func(double ch4, double n2o) {
    return 0.47 * log(1.0 + 2.01E-5 * pow(ch4 * n2o, 0.75) + 5.31E-15 * ch4 * pow(ch4 * n2o, 1.52));
}

// slot_only: form the surface slot again from the last call's outputs and
// the current sea-ice fluxes, leaving everything else as it is
__global__ __launch_bounds__(SFX_NT) void k_sfx(const MemberDev* __restrict__ mems, const int32_t* __restrict__ list,
                                                 int slot_only) {
    extern __shared__ double s_pptn[];
    __shared__ double s_mean;
    const MemberDev& M = mems[list[blockIdx.x]];
    const SfxGridDev& S = *M.sg;
    const GridDev& g = M.g;
    const int mi = g.maxi, mj = g.maxj, mk = g.maxk, ncol = g.ncol;
    const size_t nc = (size_t)ncol;
    AS_GLOBAL double* sw = gp(M.sw);
    const AS_GLOBAL double* atemp_a = sw + SW_ATEMP * nc;
    AS_GLOBAL double* ashum_a = sw + SW_ASHUM * nc;
    const AS_GLOBAL double* sich_a = sw + SW_SICH * nc;
    const AS_GLOBAL double* sica_a = sw + SW_SICA * nc;
    AS_GLOBAL double* tice_a = sw + SW_TICE * nc;
    const AS_GLOBAL double* us_a = sw + SW_USURF * nc;
    AS_GLOBAL double* out = sw + SW_NF * nc;
    AS_GLOBAL double* sc = out + SF_N * nc;
    const double* solfor = S.solfor;
    // the call count istep_ocn (genie.f90:276) picks the insolation of the
    // step of the year (embm.f90:3021)
    const double istep = sc[SS_ISTEP];
    const long long istot = (long long)istep + 1;
    const int imth = (int)((istot - 1) % S.nyear) + 1;
    const AS_GLOBAL double* sol = gp(solfor) + (size_t)mj * (imth - 1);
    // greenhouse gases (2853-2926): option 4 compounds every call, option 1
    // reads CO2 from the atmosphere and, as CH4 and N2O are not carried,
    // their references; every cell holds the same value
    double co2 = sc[SS_CO2], ch4 = sc[SS_CH4], n2o = sc[SS_N2O];
    if (!S.radfor_atchem) {
        co2 = (1.0 + S.rate_co2) * co2;
        ch4 = (1.0 + S.rate_ch4) * ch4;
        n2o = (1.0 + S.rate_n2o) * n2o;
    } else {
        ch4 = CH40;
        n2o = N2O0;
    }
    // meantemp (2953): a serial sum in the reference's order, needed only
    // where the OLR's sensitivity adjustment is on; with olr_adj = 0 its
    // term is zero whatever the mean
    if (threadIdx.x == 0) {
        double mt = S.t_eqm;
        if (S.olr_adj != 0.0) {
            double a = 0.0;
            for (int c = 0; c < ncol; c++) a = a + atemp_a[c];
            mt = a / (double)(mj * mi);
        }
        s_mean = mt;
    }
    __syncthreads();
    const double meantemp = s_mean;
    const double ch4_term = ALPHACH4 * (sqrt(1.0E9 * ch4) - sqrt(1.0E9 * CH40)) - ch4_func(1.0E9 * ch4, 1.0E9 * N2O0) +
                            ch4_func(1.0E9 * CH40, 1.0E9 * N2O0);
    const double n2o_term = ALPHAN2O * (sqrt(1.0E9 * n2o) - sqrt(1.0E9 * N2O0)) - ch4_func(1.0E9 * CH40, 1.0E9 * n2o) +
                            ch4_func(1.0E9 * CH40, 1.0E9 * N2O0);
    const AS_GLOBAL double* pco2 = S.radfor_atchem ? gp(M.atm) : nullptr;

    for (int col = threadIdx.x; col < ncol && !slot_only; col += SFX_NT) {
        const int I = col % mi + 1, J = col / mi + 1;
        const double atemp = atemp_a[col], sica = sica_a[col], sich = sich_a[col];
        const double us = us_a[col], ca = gp(S.ca)[col], albcl = gp(S.albcl)[col];
        double ashum = ashum_a[col];
        double cco2 = co2;
        if (pco2) {
            const double pc = pco2[col];
            cco2 = pc < CO20 / 1000.0 ? CO20 : pc;
        }
        // 2973-3006
        const double qsata = CONST1 * exp(CONST4 * atemp / (atemp + CONST5));
        const double pptn = fmx(0.0, (ashum - S.rmax * qsata) * RHOAO * HATMBL2 * S.rdtdim);
        ashum = fmn(ashum, S.rmax * qsata);
        ashum_a[col] = ashum;
        const double rq = ashum / qsata;
        // 3020-3024
        const double fxsw = sol[J - 1] * (1.0 - albcl);
        // 3047-3077
        const double tv0 = B00 + rq * (B10 + B20 * rq);
        const double tv1 = B01 + rq * (B11 + B21 * rq);
        const double tv2 = B02 + rq * (B12 + B22 * rq);
        const double tv3 = B03 + rq * (B13 + B23 * rq);
        const double fxplw = tv0 + atemp * (tv1 + atemp * (tv2 + atemp * tv3)) - S.delf2x * log(cco2 / CO20) -
                             ch4_term - n2o_term + S.olr_adj * (meantemp - S.t_eqm) - S.olr_adj0;
        const double fxlata = RHO0 * pptn * HLV;  // 3083
        s_pptn[col] = pptn;
        double o_fxlho = 0.0, o_fxsho = 0.0, o_fxswo = 0.0, o_fxlwo = 0.0, o_fxlha, o_fxsha, o_fxswa, o_fxlwa;
        double o_dthsic = 0.0, o_dtarea = 0.0, o_albedo = albcl, o_albice = 0.0, o_fx0a, o_fx0o = 0.0;
        double o_fxsen = 0.0, o_fxlw = 0.0, o_fx0sic = 0.0, o_fx0neto = 0.0, o_tsfreez = 0.0, o_rland = 0.0;
        double evap = 0.0, evapsic = 0.0;
        if (k1at(g, I, J) <= mk) {
            const double ot = sw[SW_TSTAR * nc + col];
            // 3088-3103
            double alw = atemp + ZEROC;
            alw = alw * alw;
            alw = alw * alw;
            alw = EMA * alw;
            const double salt = SALN0 + sw[SW_SSTAR * nc + col];
            const double tsfreez = salt * (-0.0575 + 0.0017 * sqrt(salt) - 0.0002 * salt);
            const double qb = S.rsictscsf * (tsfreez - ot);
            double qbsic = qb;
            double albsic = 0.0, fx0sica = 0.0, dhsic = 0.0;
            double atm_latenti = 0.0, atm_sensiblei = 0.0, atm_netsoli = 0.0, atm_netlongi = 0.0;
            double tice = tice_a[col];
            if (sica > 0.0) {
                // 3106-3227: the ice surface temperature by Newton
                albsic = fmx(S.albsic_min, fmn(S.albsic_max, 0.40 - 0.04 * atemp));
                const double fxswsic = sol[J - 1] * (1.0 - albsic);
                double cesic, chsic, cfxsensic, qsatsic;
                for (int iter = 1; iter <= ITICE; iter++) {
                    const double ticold = tice;
                    cesic = 1.0E-3 * (1.0022 - 0.0822 * (atemp - ticold) + 0.0266 * us);
                    cesic = fmx(6.0E-5, fmn(2.19E-3, cesic));
                    chsic = 0.94 * cesic;
                    cfxsensic = RHOAIR * chsic * CPA * us;
                    qsatsic = CONST1 * exp(CONST2 * ticold / (ticold + CONST3));
                    evapsic = fmx(0.0, (qsatsic - ashum) * RHOAO * cesic * us);
                    const double tk = ticold + ZEROC, tc3 = ticold + CONST3;
                    const double tieqn = sich * ((1 - ca) * fxswsic + alw - EMO * p4(tk) - cfxsensic * (ticold - atemp) -
                                                 RHO0 * HLS * evapsic) +
                                         CONSIC * (tsfreez - ticold);
                    const double dtieq = sich * (-4.0 * EMO * p3(tk) - cfxsensic -
                                                 HLS * RHOAIR * cesic * us * qsatsic * CONST2 * CONST3 / (tc3 * tc3) * 0.5 *
                                                     (1.0 + copysign(1.0, qsatsic - ashum))) -
                                         CONSIC;
                    tice = ticold - tieqn / dtieq;
                    if (fabs(tice - ticold) < TOL || (ticold > TFREEZ && tieqn > 0.0)) break;
                }
                tice = fmn(TFREEZ, tice);
                const double fxlwsic = EMO * p4(tice + ZEROC) - alw;
                cesic = 1.0E-3 * (1.0022 - 0.0822 * (atemp - tice) + 0.0266 * us);
                cesic = fmx(6.0E-5, fmn(2.19E-3, cesic));
                chsic = 0.94 * cesic;
                cfxsensic = RHOAIR * chsic * CPA * us;
                const double fxsensic = cfxsensic * (tice - atemp);
                qsatsic = CONST1 * exp(CONST2 * tice / (tice + CONST3));
                evapsic = fmx(0.0, (qsatsic - ashum) * RHOAO * cesic * us);
                const double fx0sic = (1 - ca) * fxswsic - fxsensic - fxlwsic - RHO0 * HLS * evapsic;
                o_fx0sic = fx0sic;
                fx0sica = ca * fxswsic + fxlata + fxsensic + fxlwsic - fxplw;
                atm_latenti = +fxlata;
                atm_sensiblei = +fxsensic;
                atm_netsoli = +ca * fxswsic;
                atm_netlongi = +fxlwsic - fxplw;
                dhsic = RRHOLF * (qb - fx0sic) - RHOOI * evapsic;
                if (sich >= S.sich_max && dhsic > 0.0) {
                    qbsic = (0.0 + RHOOI * evapsic) / RRHOLF + fx0sic;
                    dhsic = RRHOLF * (qbsic - fx0sic) - RHOOI * evapsic;
                }
            } else {
                tice = 0.0;
            }
            tice_a[col] = tice;
            // 3239-3254
            const double fxlw = EMO * p4(ot + ZEROC) - alw;
            double ce = 1.0E-3 * (1.0022 - 0.0822 * (atemp - ot) + 0.0266 * us);
            ce = fmx(6.0E-5, fmn(2.19E-3, ce));
            const double ch = 0.94 * ce;
            const double fxsen = RHOAIR * ch * CPA * us * (ot - atemp);
            const double qsato = CONST1 * exp(CONST4 * ot / (ot + CONST5));
            evap = fmx(0.0, (qsato - ashum) * RHOAO * ce * us);
            // 3256-3331
            const double fx0oa = ca * fxsw + fxlata + fxsen + fxlw - fxplw;
            const double atm_latent = +fxlata, atm_sensible = +fxsen, atm_netsol = +ca * fxsw, atm_netlong = +fxlw - fxplw;
            o_fx0a = (1 - sica) * fx0oa + sica * fx0sica;
            o_fxlha = (sica * atm_latenti) + ((1 - sica) * atm_latent);
            o_fxsha = (sica * atm_sensiblei) + ((1 - sica) * atm_sensible);
            o_fxswa = (sica * atm_netsoli) + ((1 - sica) * atm_netsol);
            o_fxlwa = (sica * atm_netlongi) + ((1 - sica) * atm_netlong);
            const double fx0o = (1 - ca) * fxsw - fxsen - fxlw - RHO0 * HLV * evap;
            o_fx0neto = sica * qbsic + (1 - sica) * fmx(qb, fx0o);
            o_fxlho = (1 - sica) * (-(RHO0 * HLV * evap) + fmx(0.0, qb - fx0o)) + sica * qbsic;
            o_fxsho = -((1 - sica) * fxsen);
            o_fxswo = (1 - sica) * (1 - ca) * fxsw;
            o_fxlwo = -((1 - sica) * fxlw);
            const double dho = fmx(0.0, RRHOLF * (qb - fx0o));
            o_dthsic = sica * dhsic + (1 - sica) * dho;
            o_dtarea = fmx(0.0, RHMIN * dho * (1 - sica));
            if (sich > 1.0E-12) o_dtarea = o_dtarea + fmn(0.0, 0.5 * sica * sica * dhsic / sich);
            o_albedo = sica * albsic + (1 - sica) * albcl;
            o_albice = albsic;
            o_fx0o = fx0o;
            o_fxsen = fxsen;
            o_fxlw = fxlw;
            o_tsfreez = tsfreez;
        } else {
            // 3332-3360 (the runoff itself is gathered at its destination below)
            o_fx0a = fxsw + fxlata - fxplw;
            o_fxlha = +fxlata;
            o_fxsha = +0.0;
            o_fxswa = +fxsw;
            o_fxlwa = -fxplw;
            o_rland = pptn;
        }
        // 3640-3679
        double pptn_ocn = 0.0, evap_atm = 0.0;
        if (k1at(g, I, J) <= mk) {
            pptn_ocn = pptn;
            evap_atm = evap * (1 - sica) + evapsic * sica;
            pptn_ocn = pptn_ocn + gp(S.pmeadj)[col];
        }
        const double evap_ocn = -evap_atm;
        auto put = [&](int f, double v) { out[(size_t)f * nc + col] = v; };
        put(SF_FXLHO, o_fxlho);
        put(SF_FXSHO, o_fxsho);
        put(SF_FXSWO, o_fxswo);
        put(SF_FXLWO, o_fxlwo);
        put(SF_EVAP_OCN, evap_ocn * M2MM);
        put(SF_PPTN_OCN, pptn_ocn * M2MM);
        put(SF_RUNOFF_LAND, o_rland * M2MM);
        put(SF_FXLHA, o_fxlha);
        put(SF_FXSHA, o_fxsha);
        put(SF_FXSWA, o_fxswa);
        put(SF_FXLWA, o_fxlwa);
        put(SF_EVAP_ATM, evap_atm * M2MM);
        put(SF_PPTN_ATM, pptn * M2MM);
        put(SF_DTHSIC, o_dthsic);
        put(SF_DTAREASIC, o_dtarea);
        put(SF_ALBEDO, o_albedo);
        put(SF_ALBICE, o_albice);
        put(SF_FXSW, fxsw);
        put(SF_FX0A, o_fx0a);
        put(SF_FX0O, o_fx0o);
        put(SF_FXSEN, o_fxsen);
        put(SF_FXLW, o_fxlw);
        put(SF_EVAP, evap);
        put(SF_PPTN, pptn);
        put(SF_FX0SIC, o_fx0sic);
        put(SF_FX0NETO, o_fx0neto);
        put(SF_TSFREEZ, o_tsfreez);
    }
    __syncthreads();
    // runoff at the wet columns, then the surface slot (goldstein.f90:149-170;
    // no freshwater hosing or anomaly file)
    const AS_GLOBAL double* mjt = gp(g.mj);
    const int JS = mj + 2;
    AS_GLOBAL double* slot = gp(M.sfw);
    for (int col = threadIdx.x; col < ncol; col += SFX_NT) {
        const int J = col / mi + 1;
        double ro = 0.0;
        if (slot_only) {
            ro = out[(size_t)SF_RUNOFF_OCN * nc + col];
        } else if (k1at(g, col % mi + 1, J) <= mk) {
            double r = 0.0;
            const int s0 = gp(S.rstart)[col], s1 = gp(S.rstart)[col + 1];
            for (int q = s0; q < s1; q++) {
                const int src = gp(S.rsrc)[q];
                if (S.igrid != 0)
                    r = r + s_pptn[src] * mjt[MJ_DS * JS + src / mi + 1] * mjt[MJ_RDS * JS + J];
                else
                    r = r + s_pptn[src];
            }
            ro = (r + 0.0) * M2MM;
        }
        if (!slot_only) out[(size_t)SF_RUNOFF_OCN * nc + col] = ro;
        const double fx0neto = out[(size_t)SF_FXSWO * nc + col] + out[(size_t)SF_FXSHO * nc + col] +
                               out[(size_t)SF_FXLWO * nc + col] + out[(size_t)SF_FXLHO * nc + col] + sw[SW_FX0SIC * nc + col];
        double fw = out[(size_t)SF_PPTN_OCN * nc + col] + out[(size_t)SF_EVAP_OCN * nc + col] + ro +
                    sw[SW_FWSIC * nc + col] + 0.0 + 0.0;
        fw = fw * MM2M;
        slot[col] = -fx0neto * RFLUXSC;
        slot[nc + col] = fw * RPMESCO;
    }
    __syncthreads();
    if (threadIdx.x == 0 && !slot_only) {
        sc[SS_ISTEP] = istep + 1.0;
        sc[SS_CO2] = co2;
        sc[SS_CH4] = S.radfor_atchem ? sc[SS_CH4] : ch4;
        sc[SS_N2O] = S.radfor_atchem ? sc[SS_N2O] : n2o;
    }
}

}  // namespace

// LDS of one member's workgroup: the precipitation of every column
size_t sfx_lds(int ncol) { return (size_t)ncol * 8; }

void launch_sfx(const MemberDev* t, const int32_t* l, int n, int maxcol, int slot_only, hipStream_t s) {
    if (n <= 0) return;
    hipLaunchKernelGGL(k_sfx, dim3(n), dim3(SFX_NT), sfx_lds(maxcol), s, t, l, slot_only);
}

}  // namespace cgo
