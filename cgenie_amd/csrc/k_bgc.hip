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
// k_couple_sums    one thread per (member, wet column, tracer): the column
//             partial sums of the state after transport (biogem.f90:1984-1997,
//             2016-2025).
// k_couple_reduce  one block per member: the global sums, in the reference's
//             column order (biogem.f90:1927-2031), staged through LDS.
// k_couple_apply   one block per (member, 64 wet columns): the rescale and
//             re-normalisation of every tracer (2033-2061), one thread per
//             (column, tracer), then T, S, masses and particulates per cell.
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
__device__ __forceinline__ void stream_step(const BgcParams& p, const double* rt, int KS, int kk, double* t,
                                            double* prem) {
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
// stream spans W levels or more), kept in slot (k + off) % W; the march is
// unrolled by W so every slot index is a compile-time constant and the
// streams stay in registers.  The contributions to a level therefore arrive
// in the reference's order: DOM first, then sources from the top down, then
// the sediment return, then uptake.
template <int W>
struct Streams {
    double t[W][NSED];
    int mk[W];  // settling level of the stream, -2 for a free slot
};

template <typename R, int W>
__global__ __launch_bounds__(64) void k_bgc_col(const MemberDev* __restrict__ mems, const int32_t* __restrict__ list) {
    const MemberDev& M = mems[list[blockIdx.y]];
    const GridDev& g = M.g;
    const int w = blockIdx.x * 64 + threadIdx.x;
    if (w >= g.nwet) return;
    const BgcParams& p = M.bp;
    const int col = g.wetsorted[w];
    const int maxi = g.maxi, maxk = g.maxk;
    const int kb = k1at(g, col % maxi + 1, col / maxi + 1);
    const size_t plane = (size_t)g.ncol, ncell = (size_t)g.ncell;
    const int KS = maxk + 2;
    const double* __restrict__ Dbot = g.mk + MK_DBOT * KS;
    const double* __restrict__ Dmid = g.mk + MK_DMID * KS;
    const double* __restrict__ rt = M.rtab;
    const double* __restrict__ ocn = M.ocn;  // plane 0 = T, 1 = S, then species planes
    const double* __restrict__ Mm = M.M;
    auto cell = [&](int k) -> size_t { return (size_t)(k - 1) * plane + col; };
    const bool hasO2 = p.slot[B_O2] >= 0, hasDOMC = p.slot[B_DOMC] >= 0, hasDOMP = p.slot[B_DOMP] >= 0;
    const bool hasPO4 = p.slot[B_PO4] >= 0;

    Conv cv;
    make_conv(p, cv);

    // (1) surface carbonate chemistry (biogem.f90:1026-1084)
    double ohm = 0.0;
    if (p.slot[B_DIC] >= 0 && p.slot[B_ALK] >= 0) {
        double* carb = M.carb + col;
        const size_t cs = cell(maxk);
        const double Tsur = ocn[cs], Ssur = ocn[ncell + cs];
        double cc[NCC];
        calc_carbconst(Dmid[maxk], Tsur, Ssur, cc);
        const double Ca = salt_Ca(Ssur), B = salt_B(Ssur), SO4 = salt_SO4(Ssur), F = salt_F(Ssur);
        const double DIC = ocn[(size_t)p.ocn[B_DIC] * ncell + cs], ALK = ocn[(size_t)p.ocn[B_ALK] * ncell + cs];
        const double PO4 = hasPO4 ? ocn[(size_t)p.ocn[B_PO4] * ncell + cs] : 0.0;
        const int it = calc_carb(DIC, ALK, Ca, PO4, B, SO4, F, cc, carb, plane, p.carb_tol, p.carb_maxit);
        if (it < 0)
            atomicAdd(&M.diag[1], 1);
        else
            atomicAdd(&M.diag[2], it);
        calc_carb_RF0(DIC, ALK, PO4, B, SO4, F, cc, carb, plane, p.carb_tol, p.carb_maxit);
        ohm = carb[IC_OHM_CAL * plane];
    }

    // (2) PO4 uptake of this column (biogem_box.f90:346-1515): the same
    //     particulate and dissolved change at every level from kstart up
    int kstart = maxk + 1;
    double upart[NSED] = {0.0, 0.0, 0.0, 0.0, 0.0}, upt[4] = {0.0, 0.0, 0.0, 0.0}, domC = 0.0, domP = 0.0;
    if (hasPO4) {
        int k_mld = kb;
        const double mld = M.mld[col];
        for (int k = maxk; k >= 1; k--)
            if (Dbot[k] >= mld) {
                k_mld = k;
                break;
            }
        kstart = k_mld < kb ? kb : k_mld;
        const double lPO4 = ocn[(size_t)p.ocn[B_PO4] * ncell + cell(maxk)];
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
    double* __restrict__ part = M.part;
    double* __restrict__ vd = M.vd;
    const double* __restrict__ pO2 = hasO2 ? ocn + (size_t)p.ocn[B_O2] * ncell : nullptr;
    const double* __restrict__ pDC = hasDOMC ? ocn + (size_t)p.ocn[B_DOMC] * ncell : nullptr;
    const double* __restrict__ pDP = hasDOMP ? ocn + (size_t)p.ocn[B_DOMP] * ncell : nullptr;
    const int klim = (p.dtyr * p.sinkrate <= p.dsc) ? kb : maxk;
    const double ratio = (p.DOMlifetime > p.dtyr) ? p.dtyr / p.DOMlifetime : 1.0;
    const bool pfirst = hasDOMP && p.slot[B_DOMP] < p.slot[B_DOMC];
    const double Mkb = Mm[cell(kb)], rMkb = M.rM[cell(kb)];
    double sprev[NSED];
#pragma unroll
    for (int s = 0; s < NSED; s++) sprev[s] = M.settle[(size_t)s * plane + col];
    double settle[NSED] = {0.0, 0.0, 0.0, 0.0, 0.0};
    auto load = [&](int k, Lvl& v) {
        const size_t c = cell(k);
        v.O2 = hasO2 ? pO2[c] : 0.0;
        v.domc = hasDOMC ? pDC[c] : 0.0;
        v.domp = (hasDOMC && hasDOMP) ? pDP[c] : 0.0;
#pragma unroll
        for (int s = 0; s < NSED; s++) v.p[s] = k >= klim ? part[(size_t)s * ncell + c] : 0.0;
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
                    push(cur.p, M.mink[kk] >= kb ? M.mink[kk] : kb - 1);
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
        Streams<W> S;
#pragma unroll
        for (int j = 0; j < W; j++) S.mk[j] = -2;
        for (int base = maxk; base >= kb; base -= W) {
#pragma unroll
            for (int r = 0; r < W; r++) {
                const int kk = base - r;
                if (kk >= kb) {
                    if (kk > kb) load(kk - 1, nxt);
                    level(
                        kk, cur,
                        [&](int k_, double rf, double* v, double* np) {
#pragma unroll
                            for (int d = W - 1; d >= 1; d--) {
                                const int j = (W - 1 - r + d) % W;  // slot of source kk + d
                                if (S.mk[j] > -2) advance(S.t[j], S.mk[j], k_, rf, v, np);
                            }
                        },
                        [&](const double* t, int mk) {
                            const int j = W - 1 - r;
#pragma unroll
                            for (int s = 0; s < NSED; s++) S.t[j][s] = t[s];
                            S.mk[j] = mk;
                        });
                    cur = nxt;
                }
            }
        }
    }
#pragma unroll
    for (int s = 0; s < NSED; s++) M.settle[(size_t)s * plane + col] = settle[s];
}

// column partial sums of the state after transport, one thread per (column,
// sum): sum 0 is S (biogem.f90:2016-2025), sum q >= 1 is tracer q + 1
// (1984-1997); levels k1..maxk in order, loads issued eight levels at a time
constexpr int SUMS_THREADS = 256;
constexpr int SUMS_COLS = 64;
constexpr int SUMS_UNROLL = 8;

template <typename R>
__global__ __launch_bounds__(SUMS_THREADS) void k_couple_sums(const MemberDev* __restrict__ mems,
                                                              const int32_t* __restrict__ list) {
    __shared__ int32_t s_col[SUMS_COLS], s_kb[SUMS_COLS];
    const MemberDev& M = mems[list[blockIdx.y]];
    const GridDev& g = M.g;
    const int w0 = blockIdx.x * SUMS_COLS;
    if (w0 >= g.nwet) return;  // uniform over the block
    const int t = threadIdx.x, maxi = g.maxi, maxk = g.maxk, L = M.L;
    if (t < SUMS_COLS) {
        const int w = w0 + t;
        const int col = w < g.nwet ? g.wetsorted[w] : 0;
        s_col[t] = col;
        s_kb[t] = w < g.nwet ? k1at(g, col % maxi + 1, col / maxi + 1) : maxk + 1;
    }
    __syncthreads();
    const int cc = t % SUMS_COLS, w = w0 + cc;
    if (w >= g.nwet) return;  // no barrier below
    const int col = s_col[cc], kb = s_kb[cc], slot = g.vpos[w];
    const size_t plane = (size_t)g.ncol, ncell = (size_t)g.ncell, tstride = (size_t)maxk * plane;
    const int par = *M.parity;  // the transport step has flipped: st[par] is the new state
    const void* st = M.st[par];
    const int16_t* lo = M.lo[par];
    const double rSold = 1.0 / M.scal[SC_SMEAN];
    const double* __restrict__ os = M.ocn + ncell;
    const double* __restrict__ V = g.V;
    const double* __restrict__ Mm = M.M;
    for (int q = t / SUMS_COLS; q < L - 1; q += SUMS_THREADS / SUMS_COLS) {
        const int l = q == 0 ? 1 : q + 1;
        double acc = 0.0;
        for (int k0 = kb; k0 <= maxk; k0 += SUMS_UNROLL) {
            double x[SUMS_UNROLL], a[SUMS_UNROLL], b[SUMS_UNROLL];
#pragma unroll
            for (int u = 0; u < SUMS_UNROLL; u++) {
                if (k0 + u <= maxk) {
                    const size_t c = (size_t)(k0 + u - 1) * plane + col;
                    x[u] = StateIO<R>::ld(st, lo, (size_t)l * tstride + c);
                    a[u] = q == 0 ? V[c] : os[c];
                    b[u] = q == 0 ? 0.0 : Mm[c];
                }
            }
#pragma unroll
            for (int u = 0; u < SUMS_UNROLL; u++) {
                if (k0 + u <= maxk) {
                    if (q == 0) {
                        const double Sn = x[u] + SALN0 + 0.0;
                        acc += Sn * a[u];
                    } else {
                        const double lv = x[u] * a[u] * rSold;
                        acc += lv * b[u];
                    }
                }
            }
        }
        M.colsum[(size_t)(q == 0 ? 0 : l) * g.nwet + slot] = acc;
    }
}

// global sums in the reference's column order (partialtot(n) summed over
// n = 1..n_vocn, biogem.f90:1927-2031), then the coupling scalars.  Rows of
// column partials are staged in LDS by the whole block; each sum is then one
// thread's sequential pass over its row.
constexpr int RED_THREADS = 256;
constexpr int RED_LDS_BYTES = 56 * 1024;
constexpr int RED_MAXL = 256;

__global__ __launch_bounds__(RED_THREADS) void k_couple_reduce(const MemberDev* __restrict__ mems,
                                                               const int32_t* __restrict__ list) {
    extern __shared__ double rows[];
    __shared__ double s_val[2 * RED_MAXL];
    const MemberDev& M = mems[list[blockIdx.x]];
    const int nwet = M.g.nwet, L = M.L;
    const int stride = nwet | 1;  // odd row stride in doubles
    int nrow = RED_LDS_BYTES / (8 * stride);
    if (nrow > 16) nrow = 16;
    const bool staged = nrow > 0;  // rows longer than the LDS budget are summed from global memory
    if (!staged) nrow = RED_THREADS;
    // sum q: 0 -> S (scaled by 1/Vtot), 1 + (l-2) -> NEW_l, L-1 + (l-2) -> OLD_l
    const int nsum = 1 + 2 * (L - 2);
    auto src = [&](int q) -> const double* {
        if (q == 0) return M.colsum;
        if (q < L - 1) return M.colsum + (size_t)(q + 1) * nwet;
        return M.colold + (size_t)(q - L + 3) * nwet;
    };
    const int t = threadIdx.x;
    for (int q0 = 0; q0 < nsum; q0 += nrow) {
        const int nr = nsum - q0 < nrow ? nsum - q0 : nrow;
        if (staged) {
            for (int r = 0; r < nr; r++) {
                const double* a = src(q0 + r);
                for (int i = t; i < nwet; i += RED_THREADS) rows[r * stride + i] = a[i];
            }
        }
        __syncthreads();
        if (t < nr) {
            const double* a = staged ? rows + t * stride : src(q0 + t);
            const bool scaled = (q0 + t) == 0;
            const double sc = M.g.rVtot;
            double tot = 0.0;
            int i = 0;
            for (; i + 8 <= nwet; i += 8) {
                double v[8];
#pragma unroll
                for (int u = 0; u < 8; u++) v[u] = a[i + u];
#pragma unroll
                for (int u = 0; u < 8; u++) tot += scaled ? v[u] * sc : v[u];
            }
            for (; i < nwet; i++) tot += scaled ? a[i] * sc : a[i];
            s_val[q0 + t] = tot;
        }
        __syncthreads();
    }
    for (int l = 2 + t; l < L; l += RED_THREADS) {
        const double nw = s_val[l - 1];
        const double rnew = fabs(nw) < NULLSMALL ? 0.0 : 1.0 / nw;
        M.scal[SC_FACT0 + l] = s_val[L - 3 + l] * rnew;
    }
    if (t == 0) {
        const double Sold = M.scal[SC_SMEAN];
        const double Snew = s_val[0];
        const double Sratio = Snew / Sold;
        M.scal[SC_SOLD] = Sold;
        M.scal[SC_SNEW] = Snew;
        M.scal[SC_SMEAN] = Snew;
        M.scal[SC_SRATIO] = Sratio;
        M.scal[SC_RSRATIO] = 1.0 / Sratio;
    }
}

// the rescale of biogem_tracercoupling (biogem.f90:2001-2061).  LDS holds the
// old (ocn) and new (state) absolute salinity of the block's cells, which the
// tracer pass reads before the cell pass overwrites both.
constexpr int APPLY_THREADS = 256;
constexpr int APPLY_COLS = 64;
constexpr int APPLY_UNROLL = 8;

template <typename R>
__global__ __launch_bounds__(APPLY_THREADS) void k_couple_apply(const MemberDev* __restrict__ mems,
                                                                const int32_t* __restrict__ list) {
    extern __shared__ double sal[];  // [2][maxk][APPLY_COLS]: ocn S (old), state S (new)
    __shared__ int32_t s_col[APPLY_COLS], s_kb[APPLY_COLS];
    const MemberDev& M = mems[list[blockIdx.y]];
    const GridDev& g = M.g;
    const int w0 = blockIdx.x * APPLY_COLS;
    if (w0 >= g.nwet) return;  // uniform over the block
    const BgcParams& p = M.bp;
    const int maxi = g.maxi, maxk = g.maxk, L = M.L;
    const size_t plane = (size_t)g.ncol, ncell = (size_t)g.ncell, tstride = (size_t)maxk * plane;
    const int par = *M.parity;
    void* st = M.st[par];
    int16_t* lo = M.lo[par];
    double* __restrict__ ocn = M.ocn;
    double* sold = sal;
    double* snew = sal + (size_t)maxk * APPLY_COLS;
    const int t = threadIdx.x;
    if (t < APPLY_COLS) {
        const int w = w0 + t;
        const int col = w < g.nwet ? g.wetsorted[w] : 0;
        s_col[t] = col;
        s_kb[t] = w < g.nwet ? k1at(g, col % maxi + 1, col / maxi + 1) : maxk + 1;
    }
    __syncthreads();
    // (a) salinities of the block's cells
    for (int idx = t; idx < APPLY_COLS * maxk; idx += APPLY_THREADS) {
        const int cc = idx % APPLY_COLS, k = idx / APPLY_COLS + 1;
        if (k < s_kb[cc]) continue;
        const size_t c = (size_t)(k - 1) * plane + s_col[cc];
        sold[idx] = ocn[ncell + c];
        snew[idx] = StateIO<R>::ld(st, lo, tstride + c) + SALN0 + 0.0;
    }
    __syncthreads();
    const double rSold = 1.0 / M.scal[SC_SOLD];
    const double Snew = M.scal[SC_SNEW], Sratio = M.scal[SC_SRATIO], rSratio = M.scal[SC_RSRATIO];
    // (b) tracers l >= 3 (2033-2046) and the OLD partials of the next coupling
    {
        const int cc = t % APPLY_COLS, wv = t / APPLY_COLS;
        const int w = w0 + cc;
        if (w < g.nwet) {
            const int col = s_col[cc], kb = s_kb[cc], slot = g.vpos[w];
            for (int l = 2 + wv; l < L; l += APPLY_THREADS / APPLY_COLS) {
                int b = -1;
                for (int q = 0; q < NBGC; q++)
                    if (p.slot[q] == l) b = q;
                const bool hasv = M.bgc && b >= 0;
                const double* __restrict__ vdl = M.vd + (size_t)(b < 0 ? 0 : b) * ncell;
                double* __restrict__ ocl = ocn + (size_t)(b < 0 ? 0 : p.ocn[b]) * ncell;
                const double fact = M.scal[SC_FACT0 + l];
                double acc = 0.0;
                for (int k0 = kb; k0 <= maxk; k0 += APPLY_UNROLL) {
                    double lv[APPLY_UNROLL], vv[APPLY_UNROLL], mn[APPLY_UNROLL];
#pragma unroll
                    for (int u = 0; u < APPLY_UNROLL; u++) {
                        const int k = k0 + u;
                        if (k <= maxk) {
                            const size_t c = (size_t)(k - 1) * plane + col;
                            lv[u] = StateIO<R>::ld(st, lo, (size_t)l * tstride + c);
                            vv[u] = hasv ? vdl[c] : 0.0;
                            mn[u] = M.M[c];
                        }
                    }
#pragma unroll
                    for (int u = 0; u < APPLY_UNROLL; u++) {
                        const int k = k0 + u;
                        if (k <= maxk) {
                            const size_t c = (size_t)(k - 1) * plane + col;
                            const int li = (k - 1) * APPLY_COLS + cc;
                            const double x = lv[u] * sold[li] * rSold;
                            double v = fact * x + vv[u];
                            v = Sratio * v;
                            if (b >= 0) ocl[c] = v;
                            StateIO<R>::st(st, lo, (size_t)l * tstride + c, (Snew / snew[li]) * v);
                            acc += v * (rSratio * mn[u]);
                        }
                    }
                }
                M.colold[(size_t)l * g.nwet + slot] = acc;
            }
        }
    }
    __syncthreads();
    // (c) T, S (2001-2013), the new cell masses (2049-2052) and particulates
    for (int idx = t; idx < APPLY_COLS * maxk; idx += APPLY_THREADS) {
        const int cc = idx % APPLY_COLS, k = idx / APPLY_COLS + 1;
        if (k < s_kb[cc]) continue;
        const size_t c = (size_t)(k - 1) * plane + s_col[cc];
        const double lvT = StateIO<R>::ld(st, lo, c) + ZEROC + 0.0;
        const double lvS = snew[idx];
        ocn[c] = lvT;
        ocn[ncell + c] = lvS;
        StateIO<R>::st(st, lo, c, lvT - ZEROC);
        StateIO<R>::st(st, lo, tstride + c, lvS - SALN0);
        M.M[c] = rSratio * M.M[c];
        M.rM[c] = Sratio * M.rM[c];
#pragma unroll
        for (int s = 0; s < NSED; s++) {
            double* pp = &M.part[(size_t)s * ncell + c];
            *pp = Sratio * (*pp + 0.0);
        }
    }
}

// ------------------------------------------------------------ launchers
size_t apply_lds_bytes(int maxk) { return (size_t)2 * maxk * APPLY_COLS * 8; }

template <typename R>
static void launch_col(int W, dim3 grid, const MemberDev* mems, const int32_t* list, hipStream_t s) {
    if (W <= 4)
        hipLaunchKernelGGL((k_bgc_col<R, 4>), grid, dim3(64), 0, s, mems, list);
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

void launch_coupling(const MemberDev* mems, const int32_t* list, int nlist, int nwetmax, int maxkmax, int prec,
                     hipStream_t s) {
    const dim3 gapp((nwetmax + APPLY_COLS - 1) / APPLY_COLS, nlist);
    const dim3 gsum((nwetmax + SUMS_COLS - 1) / SUMS_COLS, nlist);
    if (prec == 0)
        hipLaunchKernelGGL(k_couple_sums<double>, gsum, dim3(SUMS_THREADS), 0, s, mems, list);
    else
        hipLaunchKernelGGL(k_couple_sums<float>, gsum, dim3(SUMS_THREADS), 0, s, mems, list);
    hipLaunchKernelGGL(k_couple_reduce, dim3(nlist), dim3(RED_THREADS), RED_LDS_BYTES, s, mems, list);
    if (prec == 0)
        hipLaunchKernelGGL(k_couple_apply<double>, gapp, dim3(APPLY_THREADS), apply_lds_bytes(maxkmax), s, mems, list);
    else
        hipLaunchKernelGGL(k_couple_apply<float>, gapp, dim3(APPLY_THREADS), apply_lds_bytes(maxkmax), s, mems, list);
}

}  // namespace cgo
