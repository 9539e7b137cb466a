// k_bgc.hip -- BIOGEM column step and biogem_tracercoupling on the device.
//
// Built with -ffp-contract=off: operand order follows the reference so the
// FP64 results track the CPU restatement to the last bits (only the device
// exp/log/pow differ from glibc, within 1 ulp).
//
// k_bgc_col   one thread per (member, wet column): the per-column part of
//             step_biogem (biogem.f90:873-969 sediment return, DOM and
//             particulate remineralisation; 1026-1104 surface carbonate
//             chemistry; 1693 -> biogem_box.f90:346-1515 PO4 uptake) and the
//             column partial sums of biogem_tracercoupling (1952-1997, 2016-2025).
// k_couple_reduce  one block per member: the global sums, in the reference's
//             column order (1927-2031).
// k_couple_apply   one thread per (member, wet column): the rescale and
//             re-normalisation of every tracer (2033-2061).
#include "carbchem_dev.h"
#include "tstep_impl.h"

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

template <typename R>
__global__ __launch_bounds__(64) void k_bgc_col(const MemberDev* __restrict__ mems, const int32_t* __restrict__ list) {
    const MemberDev& M = mems[list[blockIdx.y]];
    const GridDev& g = M.g;
    const int w = blockIdx.x * 64 + threadIdx.x;
    if (w >= g.nwet) return;
    const BgcParams& p = M.bp;
    const int col = g.wetsorted[w];
    const int slot = g.vpos[w];
    const int maxi = g.maxi, maxk = g.maxk, L = M.L;
    const int I = col % maxi + 1, J = col / maxi + 1;
    const int kb = k1at(g, I, J);
    const size_t plane = (size_t)g.ncol, ncell = (size_t)g.ncell;
    const int KS = maxk + 2;
    const double* Dbot = g.mk + MK_DBOT * KS;
    const double* Dmid = g.mk + MK_DMID * KS;
    const double* rt = M.rtab;
    double* ocn = M.ocn;  // plane 0 = T, 1 = S, then species planes
    double* vd = M.vd;
    Conv cv;
    make_conv(p, cv);
    auto ocn_of = [&](int b, int k) -> double { return ocn[(size_t)p.ocn[b] * ncell + (size_t)(k - 1) * plane + col]; };
    auto vdp = [&](int b, int k) -> double& { return vd[(size_t)b * ncell + (size_t)(k - 1) * plane + col]; };
    // bio_remin = 0 (biogem.f90:596-599)
    for (int b = 0; b < NBGC; b++)
        if (p.slot[b] >= 0)
            for (int k = kb; k <= maxk; k++) vdp(b, k) = 0.0;

    // DOM remineralisation, sub_box_remin_DOM (biogem_box.f90:2287-2406):
    // accumulated into vd (vbio_remin), before the particulate terms
    if (p.slot[B_DOMC] >= 0) {
        const double ratio = (p.DOMlifetime > p.dtyr) ? p.dtyr / p.DOMlifetime : 1.0;
        for (int k = maxk; k >= kb; k--) {
            double bp[NSED] = {0, 0, 0, 0, 0};
            if (ocn_of(B_DOMC, k) > NULLSMALL) {
                const bool pfirst = p.slot[B_DOMP] >= 0 && p.slot[B_DOMP] < p.slot[B_DOMC];
                for (int q = 0; q < 2; q++) {
                    const int b = (q == 0) != pfirst ? B_DOMC : B_DOMP;
                    if (p.slot[b] < 0) continue;
                    const int s = b == B_DOMC ? S_POC : S_POP;
                    const double v = ocn_of(b, k);
                    bp[s] = bp[s] + 1.0 * ratio * v;
                    vdp(b, k) = vdp(b, k) - ratio * v;
                }
            }
            const double rf = redfield(p, p.slot[B_O2] >= 0 ? ocn_of(B_O2, k) : 0.0);
            for (int s = 0; s < NSED; s++) {
                if (s == S_CACO3 && !p.caco3) continue;
                for (int q = 0; q < 4; q++)
                    if (p.slot[q] >= 0 && cv.c[s][q] != 0.0) vdp(q, k) = vdp(q, k) + (rf * cv.c[s][q]) * bp[s];
            }
        }
    }

    // particulate remineralisation, sub_box_remin_part (biogem_box.f90:2412-2877)
    {
        double* part = M.part;
        auto pp = [&](int s, int k) -> double& { return part[(size_t)s * ncell + (size_t)(k - 1) * plane + col]; };
        double newp[MAXK_DEV + 2][NSED];
        double settle[NSED] = {0, 0, 0, 0, 0};  // at k1 only (the level the sediment return reads)
        for (int k = kb; k <= maxk; k++)
            for (int s = 0; s < NSED; s++) newp[k][s] = 0.0;
        const int klim = (p.dtyr * p.sinkrate <= p.dsc) ? kb : maxk;
        double POC_ratio = 0.0, CaCO3_ratio = 0.0;
        for (int k = maxk; k >= klim; k--) {
            double tot = 0.0;
            tot = tot + pp(S_POC, k);
            if (p.caco3) tot = tot + pp(S_CACO3, k);
            if (!(tot > NULLSMALL)) continue;
            int min_k = kb - 1;
            if (k != kb && M.mink[k] >= kb) min_k = M.mink[k];
            double t[NSED];
            for (int s = 0; s < NSED; s++) t[s] = pp(s, k);
            // settling through level k itself (kk = k of the settle loop)
            if (k == kb) {
                for (int s = 0; s < NSED; s++)
                    settle[s] = settle[s] + ((s == S_POCFR2 || s == S_CACO3FR2)
                                                 ? t[s]
                                                 : M.M[(size_t)(k - 1) * plane + col] * t[s]);
            }
            for (int kk = k - 1; kk >= min_k; kk--) {
                if (kk < kb) continue;
                const double layerratio = rt[RT_LAYER * KS + kk];
                double tn[NSED];
                for (int s = 0; s < NSED; s++) tn[s] = 0.0;
                if (p.caco3) {
                    const double f1 = rt[RT_CAL1 * KS + kk];
                    const double f2 = rt[RT_CAL2 * KS + kk];
                    CaCO3_ratio = 1.0 - ((1.0 - t[S_CACO3FR2]) * f1 + t[S_CACO3FR2] * f2);
                    tn[S_CACO3FR2] = t[S_CACO3FR2] > NULLSMALL ? (1.0 - f2) * t[S_CACO3FR2] / CaCO3_ratio : 0.0;
                }
                {
                    const double f1 = rt[RT_POC1 * KS + kk];
                    const double f2 = rt[RT_POC2 * KS + kk];
                    POC_ratio = 1.0 - ((1.0 - t[S_POCFR2]) * f1 + t[S_POCFR2] * f2);
                    tn[S_POCFR2] = t[S_POCFR2] > NULLSMALL ? (1.0 - f2) * t[S_POCFR2] / POC_ratio : 0.0;
                }
                tn[S_POC] = t[S_POC] * layerratio * POC_ratio;
                tn[S_POP] = t[S_POP] * layerratio * POC_ratio;
                tn[S_CACO3] = p.caco3 ? t[S_CACO3] * layerratio * CaCO3_ratio : 0.0;
                double prem[NSED];
                for (int s = 0; s < NSED; s++) prem[s] = (layerratio * t[s] - tn[s]);
                const double rf = redfield(p, p.slot[B_O2] >= 0 ? ocn_of(B_O2, kk) : 0.0);
                for (int s = 0; s < NSED; s++) {
                    if (s == S_CACO3 && !p.caco3) continue;
                    for (int q = 0; q < 4; q++)
                        if (p.slot[q] >= 0 && cv.c[s][q] != 0.0) vdp(q, kk) = vdp(q, kk) + (rf * cv.c[s][q]) * prem[s];
                }
                for (int s = 0; s < NSED; s++) t[s] = tn[s];
                if (kk == kb && kk >= min_k + 1) {
                    for (int s = 0; s < NSED; s++)
                        settle[s] = settle[s] + ((s == S_POCFR2 || s == S_CACO3FR2)
                                                     ? t[s]
                                                     : M.M[(size_t)(kk - 1) * plane + col] * t[s]);
                }
            }
            if (min_k >= kb)
                for (int s = 0; s < NSED; s++) newp[min_k][s] = newp[min_k][s] + t[s];
        }
        // sediment dissolution of the previous step's rain (closed system
        // without SEDGEM, biogem.f90:873-923, 937-942), then bio_remin += vbio_remin
        {
            const double rf = redfield(p, p.slot[B_O2] >= 0 ? ocn_of(B_O2, kb) : 0.0);
            double fsed[4] = {0, 0, 0, 0};
            for (int s = 0; s < NSED; s++) {
                if (s == S_CACO3 && !p.caco3) continue;
                for (int q = 0; q < 4; q++)
                    if (cv.c[s][q] != 0.0) fsed[q] = fsed[q] + (rf * cv.c[s][q]) * M.settle[(size_t)s * plane + col];
            }
            const double rM = M.rM[(size_t)(kb - 1) * plane + col];
            for (int q = 0; q < 4; q++)
                if (p.slot[q] >= 0) vdp(q, kb) = rM * fsed[q] + vdp(q, kb);
        }
        for (int k = kb; k <= maxk; k++)
            for (int s = 0; s < NSED; s++) pp(s, k) = newp[k][s];
        for (int s = 0; s < NSED; s++) M.settle[(size_t)s * plane + col] = settle[s];
    }

    // surface carbonate chemistry (biogem.f90:1026-1084)
    double ohm = 0.0;
    double* carb = M.carb + col;
    const double Tsur = ocn[(size_t)(maxk - 1) * plane + col];
    const double Ssur = ocn[ncell + (size_t)(maxk - 1) * plane + col];
    if (p.slot[B_DIC] >= 0 && p.slot[B_ALK] >= 0) {
        double cc[NCC];
        calc_carbconst(Dmid[maxk], Tsur, Ssur, cc);
        const double Ca = salt_Ca(Ssur), B = salt_B(Ssur), SO4 = salt_SO4(Ssur), F = salt_F(Ssur);
        const double DIC = ocn_of(B_DIC, maxk), ALK = ocn_of(B_ALK, maxk);
        const double PO4 = p.slot[B_PO4] >= 0 ? ocn_of(B_PO4, maxk) : 0.0;
        const int it = calc_carb(DIC, ALK, Ca, PO4, B, SO4, F, cc, carb, plane, p.carb_tol, p.carb_maxit);
        if (it < 0)
            atomicAdd(&M.diag[1], 1);
        else
            atomicAdd(&M.diag[2], it);
        calc_carb_RF0(DIC, ALK, PO4, B, SO4, F, cc, carb, plane, p.carb_tol, p.carb_maxit);
        ohm = carb[IC_OHM_CAL * plane];
    }

    // PO4 Michaelis-Menten uptake, 1N1T_PO4MM (biogem_box.f90:346-1515)
    if (p.slot[B_PO4] >= 0) {
        int k_mld = kb;
        for (int k = maxk; k >= 1; k--)
            if (Dbot[k] >= M.mld[col]) {
                k_mld = k;
                break;
            }
        const double lPO4 = ocn_of(B_PO4, maxk);
        const double kPO4 = lPO4 / (lPO4 + p.c0_PO4);
        const double ficefree = (1.0 - 0.0);
        const double kI = M.light[col];
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
        double* part = M.part;
        for (int k = (k_mld < kb ? kb : k_mld); k <= maxk; k++) {
            double bp[NSED];
            bp[S_POC] = p.red_POP_POC * dPO4;
            bp[S_CACO3] = p.caco3 ? rCaCO3 * bp[S_POC] : 0.0;
            bp[S_POP] = red_POC_POP * bp[S_POC];
            double upt[4] = {0, 0, 0, 0};
            for (int s = 0; s < 3; s++) {
                if (s == S_CACO3 && !p.caco3) continue;
                for (int q = 0; q < 4; q++)
                    if (cv.c[s][q] != 0.0) upt[q] = upt[q] + cv.c[s][q] * bp[s];
            }
            const double domC = 1.0 * DOMfrac * bp[S_POC];
            const double domP = 1.0 * DOMfrac * bp[S_POP];
            if (p.slot[B_DOMC] >= 0) vdp(B_DOMC, k) = vdp(B_DOMC, k) + domC;
            if (p.slot[B_DOMP] >= 0) vdp(B_DOMP, k) = vdp(B_DOMP, k) + domP;
            bp[S_POC] = bp[S_POC] - (domC + 0.0);
            bp[S_POP] = bp[S_POP] - (domP + 0.0);
            bp[S_CACO3] = bp[S_CACO3] - (0.0 + 0.0);
            bp[S_POCFR2] = fr2;
            bp[S_CACO3FR2] = p.caco3 ? p.CaCO3_frac2 : 0.0;
            for (int s = 0; s < NSED; s++) part[(size_t)s * ncell + (size_t)(k - 1) * plane + col] = bp[s];
            for (int q = 0; q < 4; q++)
                if (p.slot[q] >= 0) vdp(q, k) = vdp(q, k) - upt[q];
        }
    }

    // column partial sums for the coupling (biogem.f90:1984-1997, 2016-2025)
    {
        const int par = *M.parity;
        const void* st = M.st[par ^ 1];
        const int16_t* lo = M.lo[par ^ 1];
        const size_t tstride = (size_t)maxk * plane;
        const double rSold = 1.0 / M.scal[SC_SMEAN];
        double sS = 0.0;
        for (int k = kb; k <= maxk; k++) {
            const size_t c = (size_t)(k - 1) * plane + col;
            const double Sn = StateIO<R>::ld(st, lo, tstride + c) + SALN0 + 0.0;
            sS += Sn * g.V[c];
        }
        M.colsum[slot] = sS;
        for (int l = 2; l < L; l++) {
            double s = 0.0;
            for (int k = kb; k <= maxk; k++) {
                const size_t c = (size_t)(k - 1) * plane + col;
                const double lv = StateIO<R>::ld(st, lo, (size_t)l * tstride + c) * ocn[ncell + c] * rSold;
                s += lv * M.M[c];
            }
            M.colsum[(size_t)l * g.nwet + slot] = s;
        }
    }
}

// sequential sum in the given order, loads issued ahead of the dependent adds
__device__ __forceinline__ double seq_sum(const double* __restrict__ a, int n, double scale, bool scaled) {
    double tot = 0.0;
    int i = 0;
    for (; i + 16 <= n; i += 16) {
        double v[16];
#pragma unroll
        for (int u = 0; u < 16; u++) v[u] = a[i + u];
#pragma unroll
        for (int u = 0; u < 16; u++) tot += scaled ? v[u] * scale : v[u];
    }
    for (; i < n; i++) tot += scaled ? a[i] * scale : a[i];
    return tot;
}

// global sums in the reference's column order (partialtot(n) summed over
// n = 1..n_vocn, biogem.f90:1927-2031), then the coupling scalars
__global__ __launch_bounds__(128) void k_couple_reduce(const MemberDev* __restrict__ mems,
                                                       const int32_t* __restrict__ list) {
    const MemberDev& M = mems[list[blockIdx.x]];
    const int nwet = M.g.nwet, L = M.L;
    __shared__ double s_new[256], s_old[256];
    const int t = threadIdx.x;
    for (int q = t; q < 2 * L; q += blockDim.x) {
        const int l = q < L ? q : q - L;
        if (l == 1) continue;
        if (q < L) {
            const double v = seq_sum(M.colsum + (size_t)l * nwet, nwet, M.g.rVtot, l == 0);
            if (l < 256) s_new[l] = v;
        } else if (l >= 2) {
            const double v = seq_sum(M.colold + (size_t)l * nwet, nwet, 0.0, false);
            if (l < 256) s_old[l] = v;
        }
    }
    __syncthreads();
    for (int l = 2 + t; l < L; l += blockDim.x) {
        const double nw = s_new[l];
        const double rnew = fabs(nw) < NULLSMALL ? 0.0 : 1.0 / nw;
        M.scal[SC_FACT0 + l] = s_old[l] * rnew;
    }
    if (t == 0) {
        const double Snew = s_new[0];
        const double Sratio = Snew / M.scal[SC_SMEAN];
        M.scal[SC_SNEW] = Snew;
        M.scal[SC_SRATIO] = Sratio;
        M.scal[SC_RSRATIO] = 1.0 / Sratio;
    }
}

template <typename R>
__global__ __launch_bounds__(64) void k_couple_apply(const MemberDev* __restrict__ mems,
                                                     const int32_t* __restrict__ list) {
    const MemberDev& M = mems[list[blockIdx.y]];
    const GridDev& g = M.g;
    const int w = blockIdx.x * 64 + threadIdx.x;
    if (w >= g.nwet) return;
    const BgcParams& p = M.bp;
    const int col = g.wetsorted[w];
    const int slot = g.vpos[w];
    const int maxi = g.maxi, maxk = g.maxk, L = M.L;
    const int kb = k1at(g, col % maxi + 1, col / maxi + 1);
    const size_t plane = (size_t)g.ncol, ncell = (size_t)g.ncell, tstride = (size_t)maxk * plane;
    const int par = *M.parity;
    void* st = M.st[par ^ 1];
    int16_t* lo = M.lo[par ^ 1];
    const double Sold = M.scal[SC_SMEAN], rSold = 1.0 / Sold;
    const double Snew = M.scal[SC_SNEW], Sratio = M.scal[SC_SRATIO], rSratio = M.scal[SC_RSRATIO];
    double* ocn = M.ocn;
    double sold[MAXK_DEV + 2], snew[MAXK_DEV + 2], mnew[MAXK_DEV + 2];
    // (2) T, S (biogem.f90:2001-2013) and the new cell masses (2049-2052)
    for (int k = kb; k <= maxk; k++) {
        const size_t c = (size_t)(k - 1) * plane + col;
        const double lvT = StateIO<R>::ld(st, lo, c) + ZEROC + 0.0;
        const double lvS = StateIO<R>::ld(st, lo, tstride + c) + SALN0 + 0.0;
        sold[k] = ocn[ncell + c];
        snew[k] = lvS;
        ocn[c] = lvT;
        ocn[ncell + c] = lvS;
        StateIO<R>::st(st, lo, c, lvT - ZEROC);
        StateIO<R>::st(st, lo, tstride + c, lvS - SALN0);
        mnew[k] = rSratio * M.M[c];
    }
    // (3) tracers l >= 3 (2033-2046) and OLD partials for the next coupling
    for (int l = 2; l < L; l++) {
        int b = -1;
        for (int q = 0; q < NBGC; q++)
            if (p.slot[q] == l) b = q;
        const double fact = M.scal[SC_FACT0 + l];
        double part = 0.0;
        for (int k = kb; k <= maxk; k++) {
            const size_t c = (size_t)(k - 1) * plane + col;
            const size_t ci = (size_t)l * tstride + c;
            const double lv = StateIO<R>::ld(st, lo, ci) * sold[k] * rSold;
            const double vdv = (M.bgc && b >= 0) ? M.vd[(size_t)b * ncell + c] : 0.0;
            double v = fact * lv + vdv;
            v = Sratio * v;
            if (b >= 0) ocn[(size_t)p.ocn[b] * ncell + c] = v;
            StateIO<R>::st(st, lo, ci, (Snew / snew[k]) * v);
            part += v * mnew[k];
        }
        M.colold[(size_t)l * g.nwet + slot] = part;
    }
    for (int k = kb; k <= maxk; k++) {
        const size_t c = (size_t)(k - 1) * plane + col;
        for (int s = 0; s < NSED; s++) {
            double* pp = &M.part[(size_t)s * ncell + c];
            *pp = *pp + 0.0;
            *pp = Sratio * *pp;
        }
        M.M[c] = mnew[k];
        M.rM[c] = Sratio * M.rM[c];
    }
}

// ------------------------------------------------------------ launchers
#define GRID_COLS(nwetmax) dim3(((nwetmax) + 63) / 64, nlist)

void launch_bgc(const MemberDev* mems, const int32_t* list, int nlist, int nwetmax, int prec, hipStream_t s) {
    if (prec == 0)
        hipLaunchKernelGGL(k_bgc_col<double>, GRID_COLS(nwetmax), dim3(64), 0, s, mems, list);
    else
        hipLaunchKernelGGL(k_bgc_col<float>, GRID_COLS(nwetmax), dim3(64), 0, s, mems, list);
    hipLaunchKernelGGL(k_couple_reduce, dim3(nlist), dim3(128), 0, s, mems, list);
    if (prec == 0)
        hipLaunchKernelGGL(k_couple_apply<double>, GRID_COLS(nwetmax), dim3(64), 0, s, mems, list);
    else
        hipLaunchKernelGGL(k_couple_apply<float>, GRID_COLS(nwetmax), dim3(64), 0, s, mems, list);
}

}  // namespace cgo
