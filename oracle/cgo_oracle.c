/*
 * cgo_oracle.c -- CPU restatement of the cGENIE ocean tracer timestep.
 *
 * TEST INFRASTRUCTURE ONLY: the checker for the HIP engine.  Only tests/,
 * __graft_entry__.smoke() and bench.py's cpu_baseline leg load it; the
 * product path never does.
 *
 * Built with -O2 -ffp-contract=off so that every operation is rounded as
 * written, in the operand order of the reference expressions (gfortran
 * without -ffast-math evaluates left to right and keeps parentheses).
 * Indices are the reference's Fortran indices; the I*() macros map them onto
 * the column-major arrays with the reference lower bounds.
 */
#include "cgo_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

/* goldstein_lib.f90:48-53 (all literals are REAL*8 under -fdefault-real-8) */
static const double USC = 0.05, RSC = 6.37E6, DSC = 5.0E3, FSC = 2 * 7.2921E-5;
static const double GSC = 9.81, RH0SC = 1.0E3;
/* gem_cmn.f90:510-513, 690, 717-720 */
static const double CONV_M3_KG = 1027.649;
static const double CONST_ZEROC = 273.15;
static const double CONST_R = 83.145;
static const double NULLSMALL = +0.999999E-19;
static const double CONST_PI = 3.141592653589793;
static const double CONST_REARTH = 6.37E+06;
static const double SALN0 = 34.9; /* goldstein-defaults.nml:60 */
/* gem_cmn.f90:511-545, 694-698 */
static const double CONV_YR_HR = 24.0 * 365.25;
static const double CONV_YR_S = 3600.0 * (24.0 * 365.25);
static const double CONV_S_YR = 1.0 / (3600.0 * (24.0 * 365.25));
static const double CONV_CM_M = 1.0 / 100.0;
static const double CONV_PA_ATM = 1.0 / 1.01325E+05;
static const double CONST_R_SI = 8.3145;
static const double CONST_V = 0.022414;

/* libgcc __powidf2: what gfortran calls for REAL**INTEGER (x**k with a
 * variable integer k), binary exponentiation. */
static double powi(double x, int m)
{
    unsigned n = m < 0 ? (unsigned)(-m) : (unsigned)m;
    double y = (n % 2) ? x : 1.0;
    while (n >>= 1) {
        x = x * x;
        if (n % 2) y = y * x;
    }
    return m < 0 ? 1.0 / y : y;
}

/* ------------------------------------------------------------------------ */
/* grid: goldstein.f90:906-1075, 1406-1407, 2057-2070                        */
/* ------------------------------------------------------------------------ */
void og_grid_init(og_grid *g, int maxi, int maxj, int maxk, int nyear,
                  double yearlen, double kappa_h, double kappa_v,
                  double ssmaxsurf, double ssmaxdeep)
{
    memset(g, 0, sizeof(*g));
    g->maxi = maxi;
    g->maxj = maxj;
    g->maxk = maxk;
    const double pi = 4 * atan(1.0);
    const double rhosc = RH0SC * FSC * USC * RSC / GSC / DSC;
    const double tsc = RSC / USC;
    g->tsc = tsc;
    /* 906-916 */
    const double th0 = -pi / 2, th1 = pi / 2;
    const double s0 = sin(th0), s1 = sin(th1);
    const double phix = 2 * pi;
    g->dphi = phix / maxi;
    g->rdphi = 1.0 / g->dphi;
    /* 922-942, igrid = 0 (constant d sin(lat)) */
    g->sv[0] = s0;
    g->cv[0] = cos(th0);
    const double dscon = (s1 - s0) / maxj;
    for (int j = 1; j <= maxj; j++) {
        g->sv[j] = s0 + j * dscon;
        g->cv[j] = sqrt(1 - g->sv[j] * g->sv[j]);
        g->s[j] = g->sv[j] - 0.5 * dscon;
    }
    /* 948-960 */
    for (int j = 1; j <= maxj; j++) {
        g->ds[j] = g->sv[j] - g->sv[j - 1];
        g->rds[j] = 1.0 / g->ds[j];
        g->c[j] = sqrt(1 - g->s[j] * g->s[j]);
        g->rc[j] = 1.0 / g->c[j];
        g->rc2[j] = g->rc[j] * g->rc[j] * g->rdphi;
        if (j < maxj) {
            g->dsv[j] = g->s[j + 1] - g->s[j];
            g->rdsv[j] = 1.0 / g->dsv[j];
            g->rcv[j] = 1.0 / g->cv[j];
            g->cv2[j] = g->cv[j] * g->cv[j] * g->rdsv[j];
            if (j > 1) g->rds2[j] = 2.0 / (g->dsv[j] + g->dsv[j - 1]);
        }
    }
    /* 973-977: dt(k) = 86400*yearlen/(nyear*tsc) */
    const double tv = 86400.0 * yearlen / (nyear * tsc);
    for (int k = 0; k <= maxk + 1; k++) g->dt[k] = tv;
    g->dtyr = 1.0 / nyear; /* the BIOGEM step at conv_kocn_kbiogem = 1 */
    /* 983-1011, ez0 = 0.1 */
    const double ez0 = 0.1;
    const double z1 = ez0 * (pow(1.0 + 1 / ez0, 1.0 / maxk) - 1.0);
    double tv4 = ez0 * (pow(z1 / ez0 + 1, 0.5) - 1);
    double tv2 = 0, tv1 = 0;
    g->zro[maxk] = -tv4;
    g->zw[maxk] = tv2;
    for (int k = 1; k <= maxk; k++) {
        double tv3 = ez0 * (powi(z1 / ez0 + 1, k) - 1);
        g->dz[maxk - k + 1] = tv3 - tv2;
        tv2 = tv3;
        double tv5 = ez0 * (pow(z1 / ez0 + 1, k + 0.5) - 1);
        if (k < maxk) g->dza[maxk - k] = tv5 - tv4;
        tv4 = tv5;
        tv1 = tv1 + g->dz[maxk - k + 1];
    }
    (void)tv1;
    for (int k = maxk; k >= 1; k--) {
        if (k > 1) g->zro[k - 1] = g->zro[k] - g->dza[k - 1];
        g->zw[k - 1] = g->zw[k] - g->dz[k];
    }
    /* 1053-1060 */
    for (int k = 1; k <= maxk - 1; k++) {
        g->rdz[k] = 1.0 / g->dz[k];
        g->rdza[k] = 1.0 / g->dza[k];
    }
    g->rdz[maxk] = 1.0 / g->dz[maxk];
    g->dza[maxk] = 0.0;
    /* 1063-1070, ieos = 0 */
    g->ec[0] = -0.0559 / rhosc;
    g->ec[1] = 0.7968 / rhosc;
    g->ec[2] = -0.0063 / rhosc;
    g->ec[3] = 3.7315E-5 / rhosc;
    g->ec[4] = 0.0;
    /* 1406-1407 */
    g->diff[0] = kappa_h / (RSC * USC);
    g->diff[1] = kappa_v * RSC / (USC * DSC * DSC);
    /* 2057-2070 */
    if (ssmaxsurf - ssmaxdeep < 1.0E-7 && ssmaxsurf - ssmaxdeep > -1.0E-7) {
        for (int k = 1; k <= maxk - 1; k++) g->ssmax[k] = ssmaxdeep;
    } else {
        const double mid = 0.5 * (log(ssmaxsurf) + log(ssmaxdeep));
        const double dif = 0.5 * (log(ssmaxsurf) - log(ssmaxdeep));
        const double efold = 200 / DSC, dep0 = -300 / DSC;
        for (int k = 1; k <= maxk - 1; k++) {
            double z = (g->zw[k] - dep0) / efold;
            g->ssmax[k] = exp(mid + dif * tanh(z));
        }
    }
}

/* eos, goldstein.f90:3048-3061 (ieos = 0); t**2 and t**3 are integer powers */
double og_eos(const double *ec, double t, double s)
{
    return ec[0] * t + ec[1] * s + ec[2] * (t * t) + ec[3] * (t * t * t);
}

/* eosd, goldstein.f90:3064-3082 (ieos = 0) */
void og_eosd(const double *ec, double t1, double t2, double s1, double s2,
             double rdz, double *dzrho, double *tec)
{
    double tatw = 0.5 * (t1 + t2);
    double te = -ec[0] - ec[2] * tatw * 2 - ec[3] * tatw * tatw * 3;
    *tec = te;
    *dzrho = (ec[1] * (s2 - s1) - te * (t2 - t1)) * rdz;
}

/* ------------------------------------------------------------------------ */
/* GOLDSTEIN tracer step                                                     */
/* ------------------------------------------------------------------------ */
typedef struct {
    const og_grid *g;
    int L, NI, NJ, NK; /* halo'd extents */
    double *ts, *ts1;  /* (maxl, 0:maxi+1, 0:maxj+1, 0:maxk+1) */
    const double *u;   /* (3, 0:maxi, 0:maxj, maxk) */
    const int32_t *k1; /* (0:maxi+1, 0:maxj+1) */
    double *rho;       /* (0:maxi+1, 0:maxj+1, 0:maxk) */
} og_ctx;

#define ITS(c, l, i, j, k) \
    ((size_t)((l) - 1) + (size_t)(c)->L * ((size_t)(i) + (size_t)(c)->NI * ((size_t)(j) + (size_t)(c)->NJ * (size_t)(k))))
#define TS(c, l, i, j, k) ((c)->ts[ITS(c, l, i, j, k)])
#define TS1(c, l, i, j, k) ((c)->ts1[ITS(c, l, i, j, k)])
#define K1(c, i, j) ((c)->k1[(size_t)(i) + (size_t)(c)->NI * (size_t)(j)])
#define U(c, n, i, j, k) \
    ((c)->u[(size_t)((n) - 1) + 3 * ((size_t)(i) + (size_t)((c)->NI - 1) * ((size_t)(j) + (size_t)((c)->NJ - 1) * (size_t)((k) - 1)))])
#define RHO(c, i, j, k) ((c)->rho[(size_t)(i) + (size_t)(c)->NI * ((size_t)(j) + (size_t)(c)->NJ * (size_t)(k))])

static inline int imax(int a, int b) { return a > b ? a : b; }
static inline int imin(int a, int b) { return a < b ? a : b; }

/* tstepo_flux, goldstein.f90:2436-2642 (iediff = 0, ieos = 0) */
static void tstepo_flux(og_ctx *c, int diso, int32_t *limps_out, double *dmax_out)
{
    const og_grid *g = c->g;
    const int maxi = g->maxi, maxj = g->maxj, maxk = g->maxk, maxl = c->L;
    const double dphi = g->dphi, rdphi = g->rdphi;
    const double diff1 = g->diff[0];
    double *fe = calloc(maxl, sizeof(double)), *fw = calloc(maxl, sizeof(double));
    double *fn = calloc(maxl, sizeof(double)), *fa = calloc(maxl, sizeof(double));
    double *fwsave = calloc(maxl, sizeof(double));
    double *fs = calloc((size_t)maxl * (maxi + 1), sizeof(double));             /* fs(l,i) */
    double *fb = calloc((size_t)maxl * (maxi + 1) * (maxj + 1), sizeof(double)); /* fb(l,i,j) */
    double *dxts = calloc((size_t)maxl * 4, sizeof(double)), *dyts = calloc((size_t)maxl * 4, sizeof(double));
    double *dzts = calloc(maxl, sizeof(double));
#define FS(l, i) fs[(size_t)((l) - 1) + (size_t)maxl * (size_t)(i)]
#define FB(l, i, j) fb[(size_t)((l) - 1) + (size_t)maxl * ((size_t)(i) + (size_t)(maxi + 1) * (size_t)(j))]
#define DXTS(l, n) dxts[(size_t)((l) - 1) + (size_t)maxl * (size_t)((n) - 1)]
#define DYTS(l, n) dyts[(size_t)((l) - 1) + (size_t)maxl * (size_t)((n) - 1)]
    double tv, ups[4], pec[4];
    double tec = 0, scc = 0, dzrho = 0, rdzrho = 0, slim, tv1;
    double dxrho[5], dyrho[5];
    int32_t limps = 0;
    double diffv = g->diff[1];
    double dmax = 0;
    if (diso) {
        scc = g->ec[1];
        limps = 0;
    }
    /* fb = 0 (calloc) */
    for (int k = 1; k <= maxk; k++) {
        memset(fs, 0, sizeof(double) * (size_t)maxl * (maxi + 1));
        for (int j = 1; j <= maxj; j++) {
            /* western doorway, 2472-2487 */
            pec[1] = U(c, 1, maxi, j, k) * dphi / diff1;
            ups[1] = pec[1] / (2.0 + fabs(pec[1]));
            for (int l = 1; l <= maxl; l++) {
                if (k >= imax(K1(c, maxi, j), K1(c, 1, j))) {
                    fw[l - 1] = U(c, 1, maxi, j, k) * g->rc[j] *
                                ((1.0 - ups[1]) * TS1(c, l, 1, j, k) + (1.0 + ups[1]) * TS1(c, l, maxi, j, k)) * 0.5;
                    fw[l - 1] = fw[l - 1] - (TS1(c, l, 1, j, k) - TS1(c, l, maxi, j, k)) * g->rc2[j] * diff1;
                } else {
                    fw[l - 1] = 0;
                }
                fwsave[l - 1] = fw[l - 1];
            }
            for (int i = 1; i <= maxi; i++) {
                if (k >= K1(c, i, j) && k < maxk) {
                    og_eosd(g->ec, TS1(c, 1, i, j, k), TS1(c, 1, i, j, k + 1), TS1(c, 2, i, j, k),
                            TS1(c, 2, i, j, k + 1), g->rdza[k], &dzrho, &tec);
                    if (dzrho < -1.0E-12)
                        rdzrho = 1.0 / dzrho;
                    else
                        rdzrho = -1.0E12;
                }
                pec[1] = U(c, 1, i, j, k) * dphi / diff1;
                ups[1] = pec[1] / (2.0 + fabs(pec[1]));
                pec[2] = U(c, 2, i, j, k) * g->dsv[imin(j, maxj - 1)] / diff1;
                ups[2] = pec[2] / (2.0 + fabs(pec[2]));
                pec[3] = U(c, 3, i, j, k) * g->dza[k] / diffv;
                ups[3] = pec[3] / (2.0 + fabs(pec[3]));
                for (int l = 1; l <= maxl; l++) {
                    /* east, 2525-2536 */
                    if (i == maxi) {
                        fe[l - 1] = fwsave[l - 1];
                    } else if (k < imax(K1(c, i, j), K1(c, i + 1, j))) {
                        fe[l - 1] = 0;
                    } else {
                        fe[l - 1] = U(c, 1, i, j, k) * g->rc[j] *
                                    ((1.0 - ups[1]) * TS1(c, l, i + 1, j, k) + (1.0 + ups[1]) * TS1(c, l, i, j, k)) * 0.5;
                        fe[l - 1] = fe[l - 1] - (TS1(c, l, i + 1, j, k) - TS1(c, l, i, j, k)) * g->rc2[j] * diff1;
                    }
                    /* north, 2538-2546 */
                    if (k < imax(K1(c, i, j), K1(c, i, j + 1))) {
                        fn[l - 1] = 0;
                    } else {
                        fn[l - 1] = g->cv[j] * U(c, 2, i, j, k) *
                                    ((1.0 - ups[2]) * TS1(c, l, i, j + 1, k) + (1.0 + ups[2]) * TS1(c, l, i, j, k)) * 0.5;
                        fn[l - 1] = fn[l - 1] - g->cv2[j] * (TS1(c, l, i, j + 1, k) - TS1(c, l, i, j, k)) * diff1;
                    }
                    /* above, 2548-2558 */
                    if (k < K1(c, i, j)) {
                        fa[l - 1] = 0;
                    } else if (k == maxk) {
                        fa[l - 1] = TS(c, l, i, j, maxk + 1);
                    } else {
                        fa[l - 1] = U(c, 3, i, j, k) *
                                    ((1.0 - ups[3]) * TS1(c, l, i, j, k + 1) + (1.0 + ups[3]) * TS1(c, l, i, j, k)) * 0.5;
                        fa[l - 1] = fa[l - 1] - (TS1(c, l, i, j, k + 1) - TS1(c, l, i, j, k)) * g->rdza[k] * diffv;
                    }
                }
                if (diso) {
                    /* isoneutral vertical flux, 2561-2624 */
                    if (k >= K1(c, i, j) && k < maxk) {
                        if (dzrho < -1.0E-12) {
                            tv1 = 0.0;
                            for (int knp = 0; knp <= 1; knp++) {
                                for (int nnp = 0; nnp <= 1; nnp++) {
                                    int ina = 1 + nnp + 2 * knp;
                                    for (int l = 1; l <= maxl; l++) {
                                        if (k + knp >= K1(c, i - 1 + 2 * nnp, j)) {
                                            DXTS(l, ina) = (TS1(c, l, i + nnp, j, k + knp) - TS1(c, l, i + nnp - 1, j, k + knp)) *
                                                           g->rc[j] * rdphi;
                                        } else {
                                            DXTS(l, ina) = 0.0;
                                        }
                                        if (k + knp >= K1(c, i, j - 1 + 2 * nnp)) {
                                            DYTS(l, ina) = (TS1(c, l, i, j + nnp, k + knp) - TS1(c, l, i, j + nnp - 1, k + knp)) *
                                                           g->cv[j - 1 + nnp] * g->rdsv[j + nnp - 1];
                                        } else {
                                            DYTS(l, ina) = 0.0;
                                        }
                                    }
                                    dxrho[ina] = scc * DXTS(2, ina) - tec * DXTS(1, ina);
                                    dyrho[ina] = scc * DYTS(2, ina) - tec * DYTS(1, ina);
                                    tv1 = tv1 + dxrho[ina] * dxrho[ina] + dyrho[ina] * dyrho[ina];
                                }
                            }
                            tv1 = 0.25 * tv1 * rdzrho * rdzrho;
                            if (tv1 > g->ssmax[k]) {
                                slim = g->ssmax[k] * g->ssmax[k] / (tv1 * tv1);
                                limps = limps + 1;
                            } else {
                                slim = 1.0;
                            }
                            tv1 = tv1 * slim * diff1 * g->rdza[k];
                            tv = tv1 * g->dt[k] * g->rdza[k];
                            if (tv > dmax) dmax = tv;
                            for (int l = 1; l <= maxl; l++) {
                                dzts[l - 1] = (TS1(c, l, i, j, k + 1) - TS1(c, l, i, j, k)) * g->rdza[k];
                                tv = 0;
                                for (int ina = 1; ina <= 4; ina++) {
                                    tv = tv + (2 * dzrho * DXTS(l, ina) - dxrho[ina] * dzts[l - 1]) * dxrho[ina] +
                                         (2 * dzrho * DYTS(l, ina) - dyrho[ina] * dzts[l - 1]) * dyrho[ina];
                                }
                                tv = 0.25 * slim * diff1 * tv / (dzrho * dzrho);
                                fa[l - 1] = fa[l - 1] + tv;
                            }
                        }
                    }
                }
                /* update, 2626-2636 */
                for (int l = 1; l <= maxl; l++) {
                    if (k >= K1(c, i, j)) {
                        TS(c, l, i, j, k) = TS1(c, l, i, j, k) -
                                            g->dt[k] * ((fe[l - 1] - fw[l - 1]) * rdphi + (fn[l - 1] - FS(l, i)) * g->rds[j] +
                                                        (fa[l - 1] - FB(l, i, j)) * g->rdz[k]);
                    }
                    fw[l - 1] = fe[l - 1];
                    FS(l, i) = fn[l - 1];
                    FB(l, i, j) = fa[l - 1];
                }
                RHO(c, i, j, k) = og_eos(g->ec, TS(c, 1, i, j, k), TS(c, 2, i, j, k));
            }
        }
    }
    *limps_out = limps;
    *dmax_out = dmax;
    free(fe); free(fw); free(fn); free(fa); free(fwsave); free(fs); free(fb);
    free(dxts); free(dyts); free(dzts);
#undef FS
#undef FB
#undef DXTS
#undef DYTS
}

/* co, goldstein.f90:2657-2777 (iconv = 0, ieos = 0) */
static void co(og_ctx *c, double *cost, int32_t *mldpk)
{
    const og_grid *g = c->g;
    const int maxi = g->maxi, maxj = g->maxj, maxk = g->maxk, maxl = c->L;
    int kk[OG_MAXK + 2];
    double dzm[OG_MAXK + 2];
    double *sum = calloc(maxl, sizeof(double));
    for (int q = 0; q < OG_MAXK + 2; q++) kk[q] = 0;
    for (int j = 1; j <= maxj; j++) {
        for (int i = 1; i <= maxi; i++) {
            const int k1 = K1(c, i, j);
            if (k1 > maxk) continue;
            kk[k1 - 1] = 0;
            for (int m = k1; m <= maxk; m++) {
                kk[m] = m;
                dzm[m] = g->dz[m];
            }
            int m = maxk, lastmix = 0;
            while (kk[m - 1] > 0 || (lastmix != 0 && kk[m] != maxk)) {
                if (kk[m - 1] == 0 || RHO(c, i, j, kk[m]) < RHO(c, i, j, kk[m - 1])) {
                    if (lastmix == 0 || kk[m] == maxk)
                        m = m - 1;
                    else
                        m = m + 1;
                    lastmix = 0;
                } else {
                    lastmix = 1;
                    int n = m - 1;
                    while (kk[n - 1] > 0 && RHO(c, i, j, kk[n]) >= RHO(c, i, j, kk[n - 1])) n = n - 1;
                    for (int l = 1; l <= maxl; l++) sum[l - 1] = TS(c, l, i, j, kk[m]) * dzm[kk[m]];
                    for (int ni = 1; ni <= m - n; ni++) {
                        for (int l = 1; l <= maxl; l++)
                            sum[l - 1] = sum[l - 1] + TS(c, l, i, j, kk[m - ni]) * dzm[kk[m - ni]];
                        dzm[kk[m]] = dzm[kk[m]] + dzm[kk[m - ni]];
                    }
                    for (int l = 1; l <= maxl; l++) TS(c, l, i, j, kk[m]) = sum[l - 1] / dzm[kk[m]];
                    RHO(c, i, j, kk[m]) = og_eos(g->ec, TS(c, 1, i, j, kk[m]), TS(c, 2, i, j, kk[m]));
                    int ni = m - 1;
                    while (kk[ni + 1] > 0) {
                        kk[ni] = kk[ni - m + n];
                        ni = ni - 1;
                    }
                }
            }
            m = maxk - 1;
            for (int n = maxk - 1; n >= k1; n--) {
                if (n > kk[m]) {
                    for (int l = 1; l <= maxl; l++) TS(c, l, i, j, n) = TS(c, l, i, j, kk[m + 1]);
                    RHO(c, i, j, n) = og_eos(g->ec, TS(c, 1, i, j, n), TS(c, 2, i, j, n));
                    cost[(i - 1) + maxi * (j - 1)] += 1.0;
                } else {
                    m = m - 1;
                }
            }
            if (mldpk) mldpk[2 * ((i - 1) + maxi * (j - 1))] = kk[m + 1];
        }
    }
    free(sum);
}

void og_goldstein_step(const og_grid *g, int maxl, const int32_t *k1,
                       double *go_ts, const double *u, const double *sflux,
                       double *rho, double *cost, int32_t *mldpk, int diso,
                       int32_t *limps, double *dmax)
{
    const int maxi = g->maxi, maxj = g->maxj, maxk = g->maxk;
    og_ctx c;
    c.g = g;
    c.L = maxl;
    c.NI = maxi + 2;
    c.NJ = maxj + 2;
    c.NK = maxk + 2;
    size_t n = (size_t)maxl * c.NI * c.NJ * c.NK;
    c.ts = calloc(n, sizeof(double));
    c.ts1 = calloc(n, sizeof(double));
    c.u = u;
    c.k1 = k1;
    c.rho = rho;
    /* 167-170: surface-flux slot */
    for (int j = 1; j <= maxj; j++)
        for (int i = 1; i <= maxi; i++)
            for (int l = 1; l <= maxl; l++) {
                double f = sflux ? sflux[(l - 1) + (size_t)maxl * ((i - 1) + (size_t)maxi * (j - 1))] : 0.0;
                TS(&c, l, i, j, maxk + 1) = f;
                TS1(&c, l, i, j, maxk + 1) = f;
            }
    /* 176-177: go_ts -> ts, ts1 (go_ts1 equals go_ts on wet cells at every
     * step boundary: tstepo 2423-2431, tracercoupling 2055-2056) */
    for (int k = 1; k <= maxk; k++)
        for (int j = 1; j <= maxj; j++)
            for (int i = 1; i <= maxi; i++)
                for (int l = 1; l <= maxl; l++) {
                    double v = go_ts[(l - 1) + (size_t)maxl * ((i - 1) + (size_t)maxi * ((j - 1) + (size_t)maxj * (k - 1)))];
                    TS(&c, l, i, j, k) = v;
                    TS1(&c, l, i, j, k) = v;
                }
    /* 182-187: periodic halo of ts1 */
    for (int j = 1; j <= maxj; j++) {
        for (int k = K1(&c, 0, j); k <= maxk; k++)
            for (int l = 1; l <= maxl; l++) TS1(&c, l, 0, j, k) = TS(&c, l, maxi, j, k);
        for (int k = K1(&c, maxi + 1, j); k <= maxk; k++)
            for (int l = 1; l <= maxl; l++) TS1(&c, l, maxi + 1, j, k) = TS(&c, l, 1, j, k);
    }
    /* tstepo, 2280-2432 (imld = 0, ieos = 0) */
    tstepo_flux(&c, diso, limps, dmax);
    co(&c, cost, mldpk);
    for (int j = 1; j <= maxj; j++) {
        for (int k = K1(&c, 0, j); k <= maxk; k++) {
            RHO(&c, 0, j, k) = RHO(&c, maxi, j, k);
            for (int l = 1; l <= maxl; l++) TS1(&c, l, 0, j, k) = TS(&c, l, maxi, j, k);
        }
        for (int k = K1(&c, maxi + 1, j); k <= maxk; k++) {
            RHO(&c, maxi + 1, j, k) = RHO(&c, 1, j, k);
            for (int l = 1; l <= maxl; l++) TS1(&c, l, maxi + 1, j, k) = TS(&c, l, 1, j, k);
        }
    }
    /* 443: go_ts = ts (ts1 = ts on wet cells, 2423-2431) */
    for (int k = 1; k <= maxk; k++)
        for (int j = 1; j <= maxj; j++)
            for (int i = 1; i <= maxi; i++)
                for (int l = 1; l <= maxl; l++)
                    go_ts[(l - 1) + (size_t)maxl * ((i - 1) + (size_t)maxi * ((j - 1) + (size_t)maxj * (k - 1)))] =
                        TS(&c, l, i, j, k);
    free(c.ts);
    free(c.ts1);
}

void og_inventory(const og_grid *g, int maxl, const int32_t *k1,
                  const double *go_ts, double *inv)
{
    const int maxi = g->maxi, maxj = g->maxj, maxk = g->maxk;
    for (int l = 0; l < maxl; l++) inv[l] = 0.0;
    for (int k = 1; k <= maxk; k++)
        for (int j = 1; j <= maxj; j++)
            for (int i = 1; i <= maxi; i++) {
                if (k < k1[i + (maxi + 2) * j]) continue;
                for (int l = 1; l <= maxl; l++)
                    inv[l - 1] += go_ts[(l - 1) + (size_t)maxl * ((i - 1) + (size_t)maxi * ((j - 1) + (size_t)maxj * (k - 1)))] *
                                  g->dphi * g->ds[j] * g->dz[k];
            }
}

/* ------------------------------------------------------------------------ */
/* carbonate chemistry: gem_carbchem.f90                                     */
/* ------------------------------------------------------------------------ */
/* pressure-correction coefficients, gem_carbchem.f90:33-46 */
static const double DP_H2CO3[5] = {-2.550E+1, +1.271E-1, +0.000E+0, -3.080E+0, +8.770E-2};
static const double DP_HCO3[5] = {-1.582E+1, -2.190E-2, +0.000E+0, +1.130E+0, -1.475E-1};
static const double DP_BO3H3[5] = {-2.948E+1, +1.622E-1, +2.608E-3, -2.840E+0, +0.000E+0};
static const double DP_H2O[5] = {-2.002E+1, +1.119E-1, -1.409E-3, -5.130E+0, +7.940E-2};
static const double DP_H4SIO4[5] = {-2.948E+1, +1.622E-1, +2.608E-3, -2.840E+0, +0.000E+0};
static const double DP_HSO4[5] = {-1.803E+1, +4.660E-2, +3.160E-4, -4.530E+0, +9.000E-2};
static const double DP_HF[5] = {-9.780E+0, -9.000E-3, -9.420E-4, -3.910E+0, +5.400E-2};
static const double DP_H3PO4[5] = {-1.451E+1, +1.211E-1, -3.210E-4, -2.670E+0, +4.270E-2};
static const double DP_H2PO4[5] = {-2.312E+1, +1.758E-1, -2.647E-3, -5.150E+0, +9.000E-2};
static const double DP_HPO4[5] = {-2.657E+1, +2.020E-1, -3.042E-3, -4.080E+0, +7.140E-2};
static const double DP_CACO3CAL[5] = {-4.876E+1, +5.304E-1, +0.000E+0, -1.176E+1, +3.692E-1};
static const double DP_CACO3ARG[5] = {-4.596E+1, +5.304E-1, +0.000E+0, -1.176E+1, +3.692E-1};
static const double DP_H2S[5] = {-1.107E+1, +9.000E-3, -9.420E-4, +2.890E+0, +5.400E-2};
static const double DP_NH4[5] = {-2.643E+0, +8.890E-1, -9.050E-3, -5.030E+0, +8.140E-2};

/* fun_corr_p, 1251-1263 */
static double corr_p(double TC, double P, double rRT, const double *dp)
{
    return (-(dp[0] + dp[1] * TC + dp[2] * TC * TC) + (5.0E-4 * (dp[3] + dp[4] * TC)) * P) * P * rRT;
}
/* 785-827 */
static double calc_Cl(double S) { double v = S / 1.80655; if (v < NULLSMALL) v = NULLSMALL; return v; }
static double calc_ION(double Cl) { return Cl > NULLSMALL ? 0.00147 + 0.03592 * Cl + 0.000068 * Cl * Cl : NULLSMALL; }
static double calc_I(double S) { return S > NULLSMALL ? 19.924 * S / (1000.0 - 1.005 * S) : NULLSMALL; }
/* 834-890 */
double og_calc_Ca(double S) { double v = 0.01028 * S / 35.0; return v < NULLSMALL ? NULLSMALL : v; }
double og_calc_Btot(double S) { double v = 0.000416 * S / 35.0; return v < NULLSMALL ? NULLSMALL : v; }
double og_calc_Ftot(double S) { double v = 0.00007 * S / 35.0; return v < NULLSMALL ? NULLSMALL : v; }
double og_calc_SO4tot(double S) { double v = 0.02824 * S / 35.0; return v < NULLSMALL ? NULLSMALL : v; }

/* 896-1245 */
static double lnkB(double T, double rT, double Tln, double S, double S05, double S15, double S20)
{
    return 148.0248 + 137.194 * S05 + 1.62247 * S +
           (-8966.90 - 2890.51 * S05 - 77.942 * S + 1.726 * S15 - 0.0993 * S20) * rT +
           (-24.4344 - 25.085 * S05 - 0.2474 * S) * Tln + 0.053105 * S05 * T;
}
static double lnkW(double rT, double Tln, double S, double S05)
{
    return 148.9802 - 13847.26 * rT - 23.6521 * Tln + (-5.977 + 118.67 * rT + 1.0495 * Tln) * S05 - 0.01615 * S;
}
static double lnkSi(double rT, double Tln, double I, double I05)
{
    return 117.40 - 8904.2 * rT - 19.334 * Tln + (3.5913 - 458.79 * rT) * I05 + (-1.5998 + 188.74 * rT) * I +
           (0.07871 - 12.1652 * rT) * I * I;
}
static double lnkHF(double T, double ION05) { return 1590.2 / T - 12.641 + 1.525 * ION05; }
static double lnkP1(double T, double Tln, double S, double S05)
{
    return 115.54 - 4576.752 / T - 18.453 * Tln + (0.69171 - 106.736 / T) * S05 + (-0.01844 - 0.65643 / T) * S;
}
static double lnkP2(double T, double Tln, double S, double S05)
{
    return 172.1033 - 8814.715 / T - 27.927 * Tln + (1.3566 - 160.340 / T) * S05 + (-0.05778 + 0.37335 / T) * S;
}
static double lnkP3(double T, double S, double S05)
{
    return -18.126 - 3070.75 / T + (2.81197 + 17.27039 / T) * S05 + (-0.09984 - 44.99486 / T) * S;
}
static double lnkHSO4(double rT, double Tln, double I05, double I, double I15, double I20)
{
    return 141.328 - 4276.1 * rT - 23.093 * Tln + (324.57 - 13856.0 * rT - 47.986 * Tln) * I05 +
           (-771.54 + 35474.0 * rT + 114.723 * Tln) * I - 2698.0 * rT * I15 + 1776.0 * rT * I20;
}
static double lnkH2S(double rT, double Tln, double S, double S05)
{
    return 225.838 - 13275.3 * rT - 34.6435 * Tln + 0.3449 * S05 - 0.0274 * S;
}
static double lnkNH4(double T, double rT, double S, double S05)
{
    return -6285.33 * rT + 0.0001635 * T - 0.25444 + (0.46532 - 123.7184 * rT) * S05 + (-0.01992 + 3.17556 * rT) * S;
}
static double lnk1_Mehrbach(double rT, double Tln, double S, double S20)
{
    return log(pow(10.0, -(3670.7 * rT - 62.008 + 9.7944 * Tln - 0.0118 * S + 0.000116 * S20)));
}
static double lnk2_Mehrbach(double rT, double S, double S20)
{
    return log(pow(10.0, -(1394.7 * rT + 4.777 - 0.0184 * S + 0.000118 * S20)));
}
static double lnQ_CO2(double Tr100, double rT, double S)
{
    return -60.2409 + 93.4517 * (100 * rT) + 23.3585 * log(Tr100) +
           S * (0.023517 - 0.023656 * (Tr100) + 0.0047036 * (Tr100 * Tr100));
}
static double lnQ_O2(double Tr100, double rT, double S)
{
    return -173.9894 + 255.5907 * (100.0 * rT) + 146.4813 * log(Tr100) - 22.2040 * (Tr100) +
           S * (-0.037362 + 0.016504 * (Tr100) - 0.0020564 * (Tr100 * Tr100)) - log(1.0E6) - log(0.20946);
}
static double logkcal(double T, double rT, double Tlog, double S, double S05, double S15)
{
    return -171.9065 - 0.077993 * T + 2839.319 * rT + 71.595 * Tlog + (-0.77712 + 0.0028426 * T + 178.34 * rT) * S05 -
           0.07711 * S + 0.0041249 * S15;
}
static double logkarg(double T, double rT, double Tlog, double S, double S05, double S15)
{
    return -171.945 - 0.077993 * T + 2903.293 * rT + 71.595 * Tlog + (-0.068393 + 0.0017276 * T + 88.135 * rT) * S05 -
           0.10018 * S + 0.0059415 * S15;
}

/* sub_calc_carbconst, 59-268 (Mehrbach) */
void og_calc_carbconst(double D, double T, double S, double *cc)
{
    double lT = T, lS = S;
    if (lT < (CONST_ZEROC + 2.0)) lT = CONST_ZEROC + 2.0;
    if (lT > (CONST_ZEROC + 35.0)) lT = CONST_ZEROC + 35.0;
    if (lS < 26.0) lS = 26.0;
    if (lS > 43.0) lS = 43.0;
    const double P = D / 10.0;
    const double S05 = sqrt(lS), S15 = pow(lS, 1.5), S20 = lS * lS;
    const double Tln = log(lT), Tlog = log10(lT), rT = 1.0 / lT, Tr100 = lT / 100.0;
    const double TC = lT - CONST_ZEROC, rRT = 1.0 / (CONST_R * lT);
    const double I = calc_I(lS), I05 = sqrt(I), I15 = pow(I, 1.5), I20 = I * I;
    const double Cl = calc_Cl(S); /* note: unclamped S, gem_carbchem.f90:113 */
    const double ION = calc_ION(Cl), ION05 = sqrt(ION);
    const double m2c = log(1 - 0.001005 * lS);
    const double SO4tot = og_calc_SO4tot(lS), Ftot = og_calc_Ftot(lS);
    cc[OG_ICC_KHSO4] = exp(lnkHSO4(rT, Tln, I05, I, I15, I20) + m2c);
    const double f2t = log(1.0 + SO4tot / cc[OG_ICC_KHSO4]);
    cc[OG_ICC_KHF] = exp(lnkHF(lT, ION05) + m2c + f2t);
    const double f2s = log(1.0 + SO4tot / cc[OG_ICC_KHSO4] + Ftot / cc[OG_ICC_KHF]);
    const double t2s = -f2t + f2s;
    cc[OG_ICC_K1] = exp(lnk1_Mehrbach(rT, Tln, lS, S20) + corr_p(TC, P, rRT, DP_H2CO3));
    cc[OG_ICC_K2] = exp(lnk2_Mehrbach(rT, lS, S20) + corr_p(TC, P, rRT, DP_HCO3));
    cc[OG_ICC_K] = cc[OG_ICC_K1] / cc[OG_ICC_K2];
    cc[OG_ICC_KB] = exp(lnkB(lT, rT, Tln, lS, S05, S15, S20) + m2c + t2s + corr_p(TC, P, rRT, DP_BO3H3));
    cc[OG_ICC_KW] = exp(lnkW(rT, Tln, lS, S05) + corr_p(TC, P, rRT, DP_H2O));
    cc[OG_ICC_KSI] = exp(lnkSi(rT, Tln, I, I05) + m2c + corr_p(TC, P, rRT, DP_H4SIO4));
    cc[OG_ICC_KHF] = exp(lnkHF(lT, ION05) + m2c + f2s + corr_p(TC, P, rRT, DP_HF));
    cc[OG_ICC_KHSO4] = exp(lnkHSO4(rT, Tln, I05, I, I15, I20) + m2c + f2s + corr_p(TC, P, rRT, DP_HSO4));
    cc[OG_ICC_KP1] = exp(lnkP1(lT, Tln, lS, S05) + corr_p(TC, P, rRT, DP_H3PO4));
    cc[OG_ICC_KP2] = exp(lnkP2(lT, Tln, lS, S05) + corr_p(TC, P, rRT, DP_H2PO4));
    cc[OG_ICC_KP3] = exp(lnkP3(lT, lS, S05) + corr_p(TC, P, rRT, DP_HPO4));
    cc[OG_ICC_KH2S] = exp(lnkH2S(rT, Tln, lS, S05) + t2s + corr_p(TC, P, rRT, DP_H2S));
    cc[OG_ICC_KNH4] = exp(lnkNH4(lT, rT, lS, S05) + corr_p(TC, P, rRT, DP_NH4));
    cc[OG_ICC_KCAL] = exp(corr_p(TC, P, rRT, DP_CACO3CAL)) * pow(10.0, logkcal(lT, rT, Tlog, lS, S05, S15));
    cc[OG_ICC_KARG] = exp(corr_p(TC, P, rRT, DP_CACO3ARG)) * pow(10.0, logkarg(lT, rT, Tlog, lS, S05, S15));
    cc[OG_ICC_QCO2] = exp(lnQ_CO2(Tr100, rT, lS));
    cc[OG_ICC_QO2] = exp(lnQ_O2(Tr100, rT, lS));
}

/* one speciation sweep shared by sub_calc_carb and sub_calc_carb_RF0
 * (gem_carbchem.f90:355-427 and 577-650): returns H1, H2 and species */
typedef struct {
    double H4BO4, OH, HPO4, PO4, H3SiO4, NH3, HS, HSO4, HF, H3PO4;
    double ALK_DIC, CO2, CO3, HCO3, H1, H2;
} og_spec;

static void carb_sweep(double H, double DIC, double ALK, double PO4t, double SiO2t, double Bt, double SO4t, double Ft,
                       double H2St, double NH4t, const double *cc, og_spec *o)
{
    const double H2 = H * H, H3 = H * H2;
    double HS = 0.0, NH3 = 0.0;
    o->OH = cc[OG_ICC_KW] / H;
    o->H4BO4 = Bt / (1.0 + H / cc[OG_ICC_KB]);
    o->H3SiO4 = SiO2t / (1.0 + H / cc[OG_ICC_KSI]);
    o->HSO4 = SO4t / (1.0 + cc[OG_ICC_KHSO4] / H);
    o->HF = Ft / (1.0 + cc[OG_ICC_KHF] / H);
    if (H2St > NULLSMALL) HS = H2St / (1.0 + H / cc[OG_ICC_KH2S]);
    if (NH4t > NULLSMALL) NH3 = NH4t / (1.0 + H / cc[OG_ICC_KNH4]);
    o->HS = HS;
    o->NH3 = NH3;
    const double P1 = cc[OG_ICC_KP1], P2 = cc[OG_ICC_KP2], P3 = cc[OG_ICC_KP3];
    o->H3PO4 = PO4t / (1.0 + P1 / H + (P1 * P2) / H2 + (P1 * P2 * P3) / H3);
    /* H2PO4 is computed by the reference but feeds nothing */
    o->HPO4 = PO4t / (1.0 + H / P2 + H2 / (P1 * P2) + P3 / H);
    o->PO4 = PO4t / (1.0 + H / P3 + H2 / (P2 * P3) + H3 / (P1 * P2 * P3));
    o->ALK_DIC = ALK - o->H4BO4 - o->OH - o->HPO4 - 2.0 * o->PO4 - o->H3SiO4 - o->NH3 - o->HS + H + o->HSO4 + o->HF +
                 o->H3PO4;
    const double K = cc[OG_ICC_K], A = o->ALK_DIC;
    const double t = 4.0 * A + DIC * K - A * K;
    const double zed = sqrt(t * t + 4.0 * (K - 4.0) * (A * A)); /* x**0.5 lowers to sqrt */
    o->HCO3 = (DIC * K - zed) / (K - 4.0);
    o->CO3 = (A * K - DIC * K - 4.0 * A + zed) / (2.0 * (K - 4.0));
    o->CO2 = DIC - A + (A * K - DIC * K - 4.0 * A + zed) / (2.0 * (K - 4.0));
    o->H1 = cc[OG_ICC_K1] * o->CO2 / o->HCO3;
    o->H2 = cc[OG_ICC_K2] * o->HCO3 / o->CO3;
}

int og_calc_carb(double DIC, double ALK, double Ca, double PO4, double SiO2,
                 double Btot, double SO4tot, double Ftot, double H2S,
                 double NH4, const double *cc, double *carb, double *carbalk,
                 double tol, int maxit)
{
    int n = 1;
    double H = carb[OG_IC_H];
    og_spec o;
    for (;;) {
        const double H_old = H;
        carb_sweep(H, DIC, ALK, PO4, SiO2, Btot, SO4tot, Ftot, H2S, NH4, cc, &o);
        if ((o.H1 < NULLSMALL) || (o.H2 < NULLSMALL)) return -1; /* 431-471, ctrl_carbchem_fail */
        H = sqrt(o.H1 * o.H2);
        if (fabs(1.0 - H / H_old) < (1.0E-8 / H) * tol) {
            carb[OG_IC_CONC_CO2] = o.CO2;
            carb[OG_IC_CONC_CO3] = o.CO3;
            carb[OG_IC_CONC_HCO3] = o.HCO3;
            carb[OG_IC_FUG_CO2] = o.CO2 / cc[OG_ICC_QCO2];
            carb[OG_IC_OHM_CAL] = Ca * o.CO3 / cc[OG_ICC_KCAL];
            carb[OG_IC_OHM_ARG] = Ca * o.CO3 / cc[OG_ICC_KARG];
            carb[OG_IC_H] = H;
            const double sat_cal = cc[OG_ICC_KCAL] * 1.0 / Ca;
            const double sat_arg = cc[OG_ICC_KARG] * 1.0 / Ca;
            carb[OG_IC_DCO3_CAL] = o.CO3 - sat_cal;
            carb[OG_IC_DCO3_ARG] = o.CO3 - sat_arg;
            if (carbalk) {
                carbalk[0] = o.HCO3;
                carbalk[1] = 2.0 * o.CO2; /* as written in the reference, 498-499 */
                carbalk[2] = o.H4BO4;
                carbalk[3] = o.OH;
                carbalk[4] = o.HPO4;
                carbalk[5] = 2.0 * o.PO4;
                carbalk[6] = o.H3SiO4;
                carbalk[7] = o.NH3;
                carbalk[8] = o.HS;
                carbalk[9] = -H;
                carbalk[10] = -o.HSO4;
                carbalk[11] = -o.HF;
                carbalk[12] = -o.H3PO4;
            }
            return n;
        }
        n = n + 1;
        if (n > maxit) return -2; /* 514-530 */
    }
}

int og_calc_carb_RF0(double DIC, double ALK, double PO4, double SiO2,
                     double Btot, double SO4tot, double Ftot, double H2S,
                     double NH4, const double *cc, double *carb, double tol,
                     int maxit)
{
    int n = 1;
    double H = carb[OG_IC_H];
    const double DIC_RF0 = DIC + 1.0e-6;
    og_spec o;
    for (;;) {
        const double H_old = H;
        carb_sweep(H, DIC_RF0, ALK, PO4, SiO2, Btot, SO4tot, Ftot, H2S, NH4, cc, &o);
        H = sqrt(o.H1 * o.H2);
        if (fabs(1.0 - H / H_old) < tol) {
            carb[OG_IC_RF0] = (o.CO2 / carb[OG_IC_CONC_CO2] - 1.0) / (DIC_RF0 / DIC - 1.0);
            return n;
        }
        n = n + 1;
        if ((o.H1 < NULLSMALL) || (o.H2 < NULLSMALL) || (n > maxit)) {
            carb[OG_IC_RF0] = 0.0;
            return -1;
        }
    }
}

/* ------------------------------------------------------------------------ */
/* BIOGEM                                                                    */
/* ------------------------------------------------------------------------ */
void og_bgc_default_params(og_bgc_params *p, double dtyr)
{
    memset(p, 0, sizeof(*p));
    p->l_DIC = p->l_ALK = p->l_PO4 = p->l_O2 = p->l_DOMC = p->l_DOMP = -1;
    p->caco3 = 1;
    p->dtyr = dtyr;
    p->k0_PO4 = 2.0E-06;
    p->c0_PO4 = 0.050E-06;
    p->DOMfrac = 0.66;
    p->DOMlifetime = 0.5;
    p->red_POP_POC = 106.0;
    p->red_POP_PO2 = -138.0;
    p->red_PON_ALK = -1.00;
    p->red_POP_PON = 16.0;
    p->red_POC_CaCO3 = 0.2;
    p->red_POC_CaCO3_pP = 0.0;
    p->POC_frac2 = 0.05;
    p->POC_dfrac2 = 0.0;
    p->POC_c0frac2 = 0.1E-6;
    p->POC_eL1 = 500.0;
    p->POC_eL2 = 1000000.0;
    p->CaCO3_frac2 = 0.5;
    p->CaCO3_eL1 = 1000.0;
    p->CaCO3_eL2 = 1000000.0;
    p->sinkrate = 125.0 / (1.0 / 365.25); /* biogem_data.f90:421 */
    p->remin_k_O2 = 1.0;
    p->remin_c0_O2 = 8.0E-6;
    p->carb_tol = 0.001;
    p->carb_maxit = 100;
    p->dsc = DSC;
    p->airsea = 0;
    p->gastransfer_a = 0.310;  /* biogem-defaults.nml:221 */
    p->atm_pCO2 = 278.0E-6;    /* atchem-defaults.nml:4 */
    p->atm_pO2 = 0.2095;       /* the modern atmosphere (the namelist default is 0) */
}

#define O3(st, n, i, j, k) ((size_t)(n) + (size_t)(st)->maxl * ((size_t)((i) - 1) + (size_t)(st)->maxi * ((size_t)((j) - 1) + (size_t)(st)->maxj * (size_t)((k) - 1))))
#define P3(st, n, i, j, k) ((size_t)(n) + (size_t)OG_NSED * ((size_t)((i) - 1) + (size_t)(st)->maxi * ((size_t)((j) - 1) + (size_t)(st)->maxj * (size_t)((k) - 1))))
#define C3(st, i, j, k) ((size_t)((i) - 1) + (size_t)(st)->maxi * ((size_t)((j) - 1) + (size_t)(st)->maxj * (size_t)((k) - 1)))
#define C2(st, i, j) ((size_t)((i) - 1) + (size_t)(st)->maxi * (size_t)((j) - 1))
#define KK1(i, j) (k1[(i) + (maxi + 2) * (j)])

/* conv_sed_ocn for the selected species (gem_util.f90:79-98,
 * biogem_data.f90:751-789); conv[ls][slot] */
typedef struct { double c[OG_NSED][4]; int slot[4]; int n; } og_conv;
/* output columns of og_conv: 0 DIC, 1 PO4, 2 O2, 3 ALK (the io order 3, 8, 10, 12) */
static void bgc_conv(const og_bgc_params *p, og_conv *cv)
{
    memset(cv, 0, sizeof(*cv));
    cv->slot[0] = p->l_DIC;
    cv->slot[1] = p->l_PO4;
    cv->slot[2] = p->l_O2;
    cv->slot[3] = p->l_ALK;
    const double cO2POP = -4.0 / 2.0;
    const double cO2PON = 0.0; /* NO3 not selected */
    const double cO2POC = p->red_POP_PO2 / p->red_POP_POC - cO2POP / p->red_POP_POC - cO2PON * p->red_POP_PON / p->red_POP_POC;
    cv->c[OG_S_POC][0] = 1.0;
    cv->c[OG_S_POC][2] = cO2POC;
    cv->c[OG_S_POP][1] = 1.0;
    cv->c[OG_S_POP][2] = cO2POP;
    cv->c[OG_S_POP][3] = p->red_PON_ALK * p->red_POP_PON;
    cv->c[OG_S_CACO3][0] = 1.0;
    cv->c[OG_S_CACO3][3] = 2.0;
}

/* sub_box_remin_redfield (biogem_box.f90:2100-2203), O2 the only oxidant:
 * the factor that multiplies conv_ls_lo */
static double redfield_factor(const og_bgc_params *p, const double *ocn_col_l)
{
    if (p->l_O2 < 0) return 0.0;
    double lk = 0.0;
    double O2 = ocn_col_l[p->l_O2];
    if (O2 < 0.0) O2 = 0.0;
    double kO2 = O2 / (O2 + p->remin_c0_O2);
    lk = lk + p->remin_k_O2 * kO2;
    if (lk < NULLSMALL) lk = 1.0;
    if (O2 < NULLSMALL) kO2 = 1.0;
    return p->remin_k_O2 * kO2 / lk;
}

static double Smean_of(const og_bgc_state *st, const int32_t *k1, const double *ocn_S_src, int stride_l,
                       double rV)
{
    /* sum over wet columns (i outer, j inner) of column partial*rV */
    const int maxi = st->maxi, maxj = st->maxj, maxk = st->maxk;
    double tot = 0.0;
    for (int i = 1; i <= maxi; i++)
        for (int j = 1; j <= maxj; j++) {
            int kb = KK1(i, j);
            if (kb > maxk) continue;
            double part = 0.0;
            for (int k = kb; k <= maxk; k++)
                part += ocn_S_src[(size_t)stride_l * C3(st, i, j, k)] * st->V[C3(st, i, j, k)];
            tot += part * rV;
        }
    return tot;
}

void og_bgc_init(og_bgc_state *st, const og_grid *g, const og_bgc_params *p,
                 int maxl, const int32_t *k1, const double *go_ts,
                 const double *light, const double *mld)
{
    const int maxi = g->maxi, maxj = g->maxj, maxk = g->maxk;
    memset(st, 0, sizeof(*st));
    st->maxi = maxi;
    st->maxj = maxj;
    st->maxk = maxk;
    st->maxl = maxl;
    size_t ncell = (size_t)maxi * maxj * maxk;
    st->ocn = calloc(ncell * maxl, sizeof(double));
    st->part = calloc(ncell * OG_NSED, sizeof(double));
    st->settle = calloc(ncell * OG_NSED, sizeof(double));
    st->carb = calloc((size_t)maxi * maxj * OG_N_CARB, sizeof(double));
    st->carbconst = calloc((size_t)maxi * maxj * OG_N_CARBCONST, sizeof(double));
    st->M = calloc(ncell, sizeof(double));
    st->rM = calloc(ncell, sizeof(double));
    st->V = calloc(ncell, sizeof(double));
    st->light = calloc((size_t)maxi * maxj, sizeof(double));
    st->mld = calloc((size_t)maxi * maxj, sizeof(double));
    memcpy(st->light, light, sizeof(double) * maxi * maxj);
    memcpy(st->mld, mld, sizeof(double) * maxi * maxj);
    st->wspeed = calloc((size_t)maxi * maxj, sizeof(double));
    st->atm = calloc((size_t)2 * maxi * maxj, sizeof(double));
    st->fatm = calloc((size_t)2 * maxi * maxj, sizeof(double));
    st->sumatm = calloc((size_t)2 * maxi * maxj, sizeof(double));
    /* sub_init_tracer_atm_comp (atchem_data.f90:230-260): every cell */
    for (size_t c = 0; c < (size_t)maxi * maxj; c++) {
        st->atm[2 * c] = p->atm_pCO2;
        st->atm[2 * c + 1] = p->atm_pO2;
    }
    /* biogem_data.f90:1098-1140 */
    double dzl[OG_MAXK + 3], dzal[OG_MAXK + 3];
    for (int k = 0; k <= maxk + 1; k++) dzl[k] = dzal[k] = 0.0;
    for (int k = 1; k <= maxk; k++) {
        dzl[k] = g->dz[k];
        dzal[k] = g->dza[k];
    }
    dzal[maxk] = dzl[maxk] / 2.0;
    for (int k = 1; k <= maxk; k++) {
        double sm = 0.0, sb = 0.0;
        for (int q = k; q <= maxk; q++) sm += p->dsc * dzal[q];
        for (int q = k; q <= maxk; q++) sb += p->dsc * dzl[q];
        st->Dmid[k] = sm;
        st->Dbot[k] = sb;
        st->dD[k] = p->dsc * dzl[k];
    }
    for (int i = 1; i <= maxi; i++)
        for (int j = 1; j <= maxj; j++)
            for (int k = KK1(i, j); k <= maxk; k++) {
                double A = 2.0 * CONST_PI * (CONST_REARTH * CONST_REARTH) * (1.0 / maxi) * (g->sv[j] - g->sv[j - 1]);
                double V = st->dD[k] * A;
                double M = CONV_M3_KG * V;
                st->V[C3(st, i, j, k)] = V;
                st->M[C3(st, i, j, k)] = M;
                st->rM[C3(st, i, j, k)] = 1.0 / M;
            }
    /* ocn from the GOLDSTEIN state: ocn = ts*S/Smean (the inverse of the
     * normalisation at biogem.f90:2046), T + 273.15, S + 34.9 */
    double Vtot = 0.0;
    for (int i = 1; i <= maxi; i++)
        for (int j = 1; j <= maxj; j++) {
            int kb = KK1(i, j);
            if (kb > maxk) continue;
            double part = 0.0;
            for (int k = kb; k <= maxk; k++) part += st->V[C3(st, i, j, k)];
            Vtot += part;
        }
    double *Sabs = calloc(ncell, sizeof(double));
    for (int i = 1; i <= maxi; i++)
        for (int j = 1; j <= maxj; j++)
            for (int k = KK1(i, j); k <= maxk; k++) Sabs[C3(st, i, j, k)] = go_ts[O3(st, 1, i, j, k)] + SALN0;
    double Smean = Smean_of(st, k1, Sabs, 1, 1.0 / Vtot);
    for (int i = 1; i <= maxi; i++)
        for (int j = 1; j <= maxj; j++)
            for (int k = KK1(i, j); k <= maxk; k++) {
                st->ocn[O3(st, 0, i, j, k)] = go_ts[O3(st, 0, i, j, k)] + CONST_ZEROC;
                st->ocn[O3(st, 1, i, j, k)] = Sabs[C3(st, i, j, k)];
                for (int l = 2; l < maxl; l++)
                    st->ocn[O3(st, l, i, j, k)] = go_ts[O3(st, l, i, j, k)] * Sabs[C3(st, i, j, k)] / Smean;
            }
    free(Sabs);
    /* biogem_data.f90:2373: [H] start, then a surface solve */
    for (int i = 1; i <= maxi; i++)
        for (int j = 1; j <= maxj; j++) {
            if (KK1(i, j) > maxk) continue;
            double *carb = &st->carb[OG_N_CARB * C2(st, i, j)];
            carb[OG_IC_H] = pow(10.0, -7.8);
        }
}

void og_bgc_free(og_bgc_state *st)
{
    free(st->ocn); free(st->part); free(st->settle); free(st->carb); free(st->carbconst);
    free(st->M); free(st->rM); free(st->V); free(st->light); free(st->mld);
    free(st->wspeed); free(st->atm); free(st->fatm); free(st->sumatm);
    memset(st, 0, sizeof(*st));
}

/* sub_box_remin_part for one column (biogem_box.f90:2412-2877): fixed
 * e-folding profiles, no ballast, no opal/det/Fe/H2S.  vrem (maxl x maxk)
 * accumulates remineralised tracer, settle (OG_NSED x maxk) is rewritten. */
static void remin_part_col(og_bgc_state *st, const og_bgc_params *p, const og_conv *cv, int i, int j, int kb,
                           double *vrem)
{
    const int maxk = st->maxk, maxl = st->maxl;
    double OLD[OG_MAXK + 2][OG_NSED], TMP[OG_MAXK + 2][OG_NSED], NEWP[OG_MAXK + 2][OG_NSED];
    double SET[OG_MAXK + 2][OG_NSED];
    double POC_ratio = 0.0, CaCO3_ratio = 0.0;
    for (int k = 0; k <= maxk + 1; k++)
        for (int s = 0; s < OG_NSED; s++) OLD[k][s] = TMP[k][s] = NEWP[k][s] = SET[k][s] = 0.0;
    for (int k = kb; k <= maxk; k++)
        for (int s = 0; s < OG_NSED; s++) OLD[k][s] = st->part[P3(st, s, i, j, k)];
    const int klim = (p->dtyr * p->sinkrate <= p->dsc) ? kb : maxk;
    for (int k = maxk; k >= klim; k--) {
        double tot = 0.0;
        tot = tot + OLD[k][OG_S_POC];
        if (p->caco3) tot = tot + OLD[k][OG_S_CACO3];
        if (!(tot > NULLSMALL)) continue;
        int min_k = kb - 1;
        if (k != kb) {
            const double maxD = st->Dbot[k] + p->dtyr * p->sinkrate;
            for (int kk = k - 1; kk >= kb; kk--)
                if (st->Dbot[kk] > maxD) {
                    min_k = kk;
                    break;
                }
        }
        for (int q = 0; q <= maxk + 1; q++)
            for (int s = 0; s < OG_NSED; s++) TMP[q][s] = 0.0;
        for (int s = 0; s < OG_NSED; s++) TMP[k][s] = OLD[k][s];
        for (int kk = k - 1; kk >= min_k; kk--) {
            if (kk < kb) continue;
            const double layerratio = st->dD[kk + 1] / st->dD[kk];
            const double dD = st->dD[kk];
            if (p->caco3) {
                double f1 = (1.0 - exp(-dD / p->CaCO3_eL1));
                double f2 = (1.0 - exp(-dD / p->CaCO3_eL2));
                CaCO3_ratio = 1.0 - ((1.0 - TMP[kk + 1][OG_S_CACO3FR2]) * f1 + TMP[kk + 1][OG_S_CACO3FR2] * f2);
                if (TMP[kk + 1][OG_S_CACO3FR2] > NULLSMALL)
                    TMP[kk][OG_S_CACO3FR2] = (1.0 - f2) * TMP[kk + 1][OG_S_CACO3FR2] / CaCO3_ratio;
                else
                    TMP[kk][OG_S_CACO3FR2] = 0.0;
            }
            {
                double f1 = (1.0 - exp(-dD / p->POC_eL1));
                double f2 = (1.0 - exp(-dD / p->POC_eL2));
                POC_ratio = 1.0 - ((1.0 - TMP[kk + 1][OG_S_POCFR2]) * f1 + TMP[kk + 1][OG_S_POCFR2] * f2);
                if (TMP[kk + 1][OG_S_POCFR2] > NULLSMALL)
                    TMP[kk][OG_S_POCFR2] = (1.0 - f2) * TMP[kk + 1][OG_S_POCFR2] / POC_ratio;
                else
                    TMP[kk][OG_S_POCFR2] = 0.0;
            }
            TMP[kk][OG_S_POC] = TMP[kk + 1][OG_S_POC] * layerratio * POC_ratio;
            TMP[kk][OG_S_POP] = TMP[kk + 1][OG_S_POP] * layerratio * POC_ratio;
            if (p->caco3) TMP[kk][OG_S_CACO3] = TMP[kk + 1][OG_S_CACO3] * layerratio * CaCO3_ratio;
            double prem[OG_NSED];
            for (int s = 0; s < OG_NSED; s++) prem[s] = (layerratio * TMP[kk + 1][s] - TMP[kk][s]);
            const double rf = redfield_factor(p, &st->ocn[O3(st, 0, i, j, kk)]);
            for (int s = 0; s < OG_NSED; s++) {
                if (s == OG_S_CACO3 && !p->caco3) continue;
                for (int q = 0; q < 4; q++) {
                    if (cv->slot[q] < 0 || cv->c[s][q] == 0.0) continue;
                    double *r = &vrem[cv->slot[q] + (size_t)maxl * (kk - 1)];
                    *r = *r + (rf * cv->c[s][q]) * prem[s];
                }
            }
        }
        if (min_k >= kb)
            for (int s = 0; s < OG_NSED; s++) NEWP[min_k][s] = NEWP[min_k][s] + TMP[min_k][s];
        for (int kk = k; kk >= min_k + 1; kk--)
            for (int s = 0; s < OG_NSED; s++) {
                if (s == OG_S_POCFR2 || s == OG_S_CACO3FR2)
                    SET[kk][s] = SET[kk][s] + TMP[kk][s];
                else
                    SET[kk][s] = SET[kk][s] + st->M[C3(st, i, j, kk)] * TMP[kk][s];
            }
    }
    for (int k = kb; k <= maxk; k++)
        for (int s = 0; s < OG_NSED; s++) {
            st->part[P3(st, s, i, j, k)] = NEWP[k][s];
            st->settle[P3(st, s, i, j, k)] = SET[k][s];
        }
}

/* ---------------------------------------------------------------- air-sea */
/* Schmidt-number and Bunsen-coefficient tables for CO2 and O2
 * (gem_data.f90:68-99, 105-141) */
static const double SC_CO2[4] = {2073.1, 125.62, 3.6276, 0.043219};
static const double SC_O2[4] = {1953.4, 128.00, 3.9918, 0.050091};
static const double BUN_CO2[6] = {-60.2409, 93.4517, 23.3585, 0.023517, -0.023656, 0.0047036};
static const double BUN_O2[6] = {-58.3877, 85.8079, 23.8439, -0.034892, 0.015568, -0.0019387};
static const double AIRSEA_R_DFLUX_DEQM_MAX = 1.00; /* biogem_lib.f90:708 */
static const double ATM_TH = 7777.0;                /* atchem_lib.f90:86 */

/* fun_calc_rho, gem_util.f90:670-688 */
double og_calc_rho(double T, double S)
{
    const double TC = T - CONST_ZEROC;
    return 1000.0 + (0.7968 * S - 0.0559 * TC - 0.0063 * powi(TC, 2) + 3.7315E-05 * powi(TC, 3));
}

/* fun_calc_solconst, gem_carbchem.f90:1367-1426: CO2 in mol/(kg atm); O2
 * (the default case) from its Bunsen coefficient over rho const_V */
static double solconst(const double *b, int bunsen, double T, double S, double rho)
{
    double lT, lS;
    if (T < CONST_ZEROC + 2.0)
        lT = CONST_ZEROC + 2.0;
    else if (T > (CONST_ZEROC + 35.0))
        lT = CONST_ZEROC + 35.0;
    else
        lT = T;
    if (S < 26.0)
        lS = 26.0;
    else if (S > 43.0)
        lS = 43.0;
    else
        lS = S;
    const double rT = 1.0 / lT, Tr100 = lT / 100.0;
    const double x = exp(b[0] + b[1] * (100 * rT) + b[2] * log(Tr100) +
                         lS * (b[3] + b[4] * (Tr100) + b[5] * powi(Tr100, 2)));
    return bunsen ? x / (rho * CONST_V) : x;
}

/* sub_calc_pv, biogem_box.f90:81-117: Wanninkhof (1992) piston velocity, m/yr */
static double piston(const double *c, double T, double wspeed, double gas_a)
{
    double TC = T - CONST_ZEROC;
    if (TC < 0.0) TC = 0.0;
    if (TC > 30.0) TC = 30.0;
    const double TC2 = TC * TC, TC3 = TC2 * TC;
    const double u2 = powi(wspeed, 2);
    const double Sc = c[0] - c[1] * TC + c[2] * TC2 - c[3] * TC3;
    return CONV_CM_M * CONV_YR_HR * gas_a * u2 * pow(Sc * 1.515E-3, -0.5);
}

/* fun_calc_ocnatm_flux for one gas (biogem_box.f90:123-299): the flux is
 * limited so that one step moves at most the surface layer's disequilibrium */
static double ocnatm_flux(double pv, double A, double rho, double dD, double ocnv, double atmv, double buff,
                          double dtyr)
{
    if (ocnv < NULLSMALL) ocnv = 0.0;
    if (atmv < NULLSMALL) atmv = 0.0;
    const double lA = (1.0 - 0.0) * A;
    double fo = pv * lA * rho * ocnv;
    double fa = pv * lA * rho * atmv;
    const double deqm = dD * A * rho * buff * fabs(atmv - ocnv);
    const double dflux = dtyr * fabs(fo - fa);
    if (deqm > NULLSMALL) {
        const double r = dflux / deqm;
        if (r > AIRSEA_R_DFLUX_DEQM_MAX) {
            fo = (AIRSEA_R_DFLUX_DEQM_MAX / r) * fo;
            fa = (AIRSEA_R_DFLUX_DEQM_MAX / r) * fa;
        }
    }
    return fo - fa;
}

void og_airsea_cell(double T, double S, double rho, double dD, double A, double wspeed, double DIC, double O2,
                    double CO2, double RF0, double pCO2, double pO2, double dtyr, double gas_a, double *out)
{
    const double sCO2 = solconst(BUN_CO2, 0, T, S, rho), sO2 = solconst(BUN_O2, 1, T, S, rho);
    const double pvCO2 = piston(SC_CO2, T, wspeed, gas_a), pvO2 = piston(SC_O2, T, wspeed, gas_a);
    /* CO2: the ocean side is [CO2], buffered by the Revelle factor */
    double atmv = sCO2 * pCO2, ocnv = CO2, buff;
    if (RF0 > NULLSMALL) {
        buff = 1.0 / (RF0 * CO2 / DIC);
    } else {
        ocnv = 0.0;
        atmv = 0.0;
        buff = 1.0;
    }
    out[0] = sCO2;
    out[1] = sO2;
    out[2] = pvCO2;
    out[3] = pvO2;
    out[4] = ocnatm_flux(pvCO2, A, rho, dD, ocnv, atmv, buff, dtyr);
    out[5] = ocnatm_flux(pvO2, A, rho, dD, O2, sO2 * pO2, 1.0, dtyr);
}

void og_bgc_set_windspeed(og_bgc_state *st, const double *wspeed)
{
    memcpy(st->wspeed, wspeed, sizeof(double) * st->maxi * st->maxj);
}

/* cpl_flux_ocnatm (atchem.f90:306-320) after BIOGEM's interface update
 * (biogem.f90:1719-1727), then step_atchem (atchem.f90:63-158) with no decay,
 * CH4 or 14C sources: the fluxes enter the atmospheric cells and the
 * atmosphere mixes to one partial pressure at 273.15 K (atchem_data.f90:248) */
void og_atchem_step(og_bgc_state *st, const og_grid *g, const og_bgc_params *p)
{
    const int maxi = g->maxi, maxj = g->maxj;
    const double dts = p->dtyr * CONV_YR_S;
    for (int q = 0; q < 2; q++) {
        for (int j = 1; j <= maxj; j++)
            for (int i = 1; i <= maxi; i++) {
                const size_t c = C2(st, i, j);
                const double A = 2.0 * CONST_PI * (CONST_REARTH * CONST_REARTH) * (1.0 / maxi) * (g->sv[j] - g->sv[j - 1]);
                const double sfx = (1.0 / A) * CONV_S_YR * st->fatm[2 * c + q];   /* 0 for land columns */
                st->sumatm[2 * c + q] = st->sumatm[2 * c + q] + dts * sfx;
            }
        double tot = 0.0, totV = 0.0;
        for (int j = 1; j <= maxj; j++)
            for (int i = 1; i <= maxi; i++) {
                const size_t c = C2(st, i, j);
                const double A = 2.0 * CONST_PI * (CONST_REARTH * CONST_REARTH) * (1.0 / maxi) * (g->sv[j] - g->sv[j - 1]);
                const double V = ATM_TH * A;
                const double cam = V / (CONV_PA_ATM * CONST_R_SI * CONST_ZEROC);
                const double cma = 1.0 / cam;
                double *a = &st->atm[2 * c + q];
                *a = *a + cma * A * st->sumatm[2 * c + q] + cma * 0.0;
            }
        for (int j = 1; j <= maxj; j++)
            for (int i = 1; i <= maxi; i++) {
                const size_t c = C2(st, i, j);
                const double A = 2.0 * CONST_PI * (CONST_REARTH * CONST_REARTH) * (1.0 / maxi) * (g->sv[j] - g->sv[j - 1]);
                const double V = ATM_TH * A;
                tot += (V / (CONV_PA_ATM * CONST_R_SI * CONST_ZEROC)) * st->atm[2 * c + q];
                totV += V;
            }
        for (int j = 1; j <= maxj; j++)
            for (int i = 1; i <= maxi; i++) {
                const size_t c = C2(st, i, j);
                st->atm[2 * c + q] = (tot / totV) * CONV_PA_ATM * CONST_R_SI * CONST_ZEROC;
                st->sumatm[2 * c + q] = 0.0;
            }
    }
}

void og_biogem_step(og_bgc_state *st, const og_grid *g, const og_bgc_params *p,
                    const int32_t *k1, double *vdocn)
{
    const int maxi = g->maxi, maxj = g->maxj, maxk = g->maxk, maxl = st->maxl;
    og_conv cv;
    bgc_conv(p, &cv);
    size_t ncell = (size_t)maxi * maxj * maxk;
    /* biogem.f90:596-599: bio_remin = 0 */
    memset(vdocn, 0, sizeof(double) * ncell * maxl);
    double *vrem = calloc((size_t)maxl * (maxk + 1), sizeof(double));
    st->carb_iter = 0;
    for (int i = 1; i <= maxi; i++)
        for (int j = 1; j <= maxj; j++) {
            const int kb = KK1(i, j);
            if (kb > maxk) continue;
            /* sediment dissolution, closed system without SEDGEM
             * (biogem.f90:873-923, 937-942) */
            {
                const double rf = redfield_factor(p, &st->ocn[O3(st, 0, i, j, kb)]);
                double fsed[4] = {0, 0, 0, 0};
                for (int s = 0; s < OG_NSED; s++) {
                    if (s == OG_S_CACO3 && !p->caco3) continue;
                    for (int q = 0; q < 4; q++) {
                        if (cv.slot[q] < 0 || cv.c[s][q] == 0.0) continue;
                        fsed[q] = fsed[q] + (rf * cv.c[s][q]) * st->settle[P3(st, s, i, j, kb)];
                    }
                }
                for (int q = 0; q < 4; q++) {
                    if (cv.slot[q] < 0) continue;
                    double *r = &vdocn[O3(st, cv.slot[q], i, j, kb)];
                    *r = *r + st->rM[C3(st, i, j, kb)] * fsed[q];
                }
            }
            /* DOM remineralisation, sub_box_remin_DOM (biogem_box.f90:2287-2406) */
            memset(vrem, 0, sizeof(double) * (size_t)maxl * (maxk + 1));
            if (p->l_DOMC >= 0) {
                for (int k = maxk; k >= kb; k--) {
                    const double *oc = &st->ocn[O3(st, 0, i, j, k)];
                    double ratio = (p->DOMlifetime > p->dtyr) ? p->dtyr / p->DOMlifetime : 1.0;
                    double bp[OG_NSED] = {0, 0, 0, 0, 0};
                    if (oc[p->l_DOMC] > NULLSMALL) {
                        /* DOM_C -> POC, DOM_P -> POP in ocean-slot order */
                        int ls[2] = {p->l_DOMC, p->l_DOMP};
                        int ss[2] = {OG_S_POC, OG_S_POP};
                        int first = (p->l_DOMP >= 0 && p->l_DOMP < p->l_DOMC) ? 1 : 0;
                        for (int q = 0; q < 2; q++) {
                            int w = first ? 1 - q : q;
                            if (ls[w] < 0) continue;
                            bp[ss[w]] = bp[ss[w]] + 1.0 * ratio * oc[ls[w]];
                            double *r = &vrem[ls[w] + (size_t)maxl * (k - 1)];
                            *r = *r - ratio * oc[ls[w]];
                        }
                    }
                    const double rf = redfield_factor(p, oc);
                    for (int s = 0; s < OG_NSED; s++) {
                        if (s == OG_S_CACO3 && !p->caco3) continue;
                        for (int q = 0; q < 4; q++) {
                            if (cv.slot[q] < 0 || cv.c[s][q] == 0.0) continue;
                            double *r = &vrem[cv.slot[q] + (size_t)maxl * (k - 1)];
                            *r = *r + (rf * cv.c[s][q]) * bp[s];
                        }
                    }
                }
            }
            /* particulate remineralisation */
            remin_part_col(st, p, &cv, i, j, kb, vrem);
            for (int k = kb; k <= maxk; k++)
                for (int l = 2; l < maxl; l++) {
                    double *r = &vdocn[O3(st, l, i, j, k)];
                    *r = *r + vrem[l + (size_t)maxl * (k - 1)];
                }
        }
    /* (i,j) loop: surface carbonate chemistry and uptake (biogem.f90:985-1105, 1693) */
    for (int i = 1; i <= maxi; i++)
        for (int j = 1; j <= maxj; j++) {
            const int kb = KK1(i, j);
            if (kb > maxk) continue;
            double *oc = &st->ocn[O3(st, 0, i, j, maxk)];
            double *cc = &st->carbconst[OG_N_CARBCONST * C2(st, i, j)];
            double *carb = &st->carb[OG_N_CARB * C2(st, i, j)];
            double ohm = 0.0;
            if (p->l_DIC >= 0 && p->l_ALK >= 0) {
                og_calc_carbconst(st->Dmid[maxk], oc[0], oc[1], cc);
                const double S = oc[1];
                const double Ca = og_calc_Ca(S), B = og_calc_Btot(S), SO4 = og_calc_SO4tot(S), F = og_calc_Ftot(S);
                const double PO4 = p->l_PO4 >= 0 ? oc[p->l_PO4] : 0.0;
                int it = og_calc_carb(oc[p->l_DIC], oc[p->l_ALK], Ca, PO4, 0.0, B, SO4, F, 0.0, 0.0, cc, carb, NULL,
                                      p->carb_tol, p->carb_maxit);
                if (it < 0)
                    st->carb_fail++;
                else
                    st->carb_iter += it;
                og_calc_carb_RF0(oc[p->l_DIC], oc[p->l_ALK], PO4, 0.0, B, SO4, F, 0.0, 0.0, cc, carb, p->carb_tol,
                                 p->carb_maxit);
                ohm = carb[OG_IC_OHM_CAL];
            }
            /* air-sea exchange (biogem.f90:1044-1046, 1605-1614) with rho from
             * biogem_climate (biogem.f90:2187), which genie calls after each
             * coupling and at start-up (genie.f90:110, 389) */
            if (p->airsea) {
                const size_t c = C2(st, i, j);
                const double A = 2.0 * CONST_PI * (CONST_REARTH * CONST_REARTH) * (1.0 / maxi) * (g->sv[j] - g->sv[j - 1]);
                const double rho = og_calc_rho(oc[0], oc[1]);
                const int co2 = p->l_DIC >= 0 && p->l_ALK >= 0;
                double f[6];
                og_airsea_cell(oc[0], oc[1], rho, st->dD[maxk], A, st->wspeed[c], co2 ? oc[p->l_DIC] : 0.0,
                               p->l_O2 >= 0 ? oc[p->l_O2] : 0.0, carb[OG_IC_CONC_CO2], carb[OG_IC_RF0],
                               st->atm[2 * c], st->atm[2 * c + 1], p->dtyr, p->gastransfer_a, f);
                st->fatm[2 * c] = co2 ? f[4] : 0.0;
                st->fatm[2 * c + 1] = p->l_O2 >= 0 ? f[5] : 0.0;
            }
            if (p->l_PO4 < 0) continue;
            /* sub_calc_bio_uptake, 1N1T_PO4MM (biogem_box.f90:346-1515) */
            int k_mld = kb;
            for (int k = maxk; k >= 1; k--)
                if (st->Dbot[k] >= st->mld[C2(st, i, j)]) {
                    k_mld = k;
                    break;
                }
            const double lPO4 = oc[p->l_PO4];
            const double kPO4 = lPO4 / (lPO4 + p->c0_PO4);
            const double ficefree = (1.0 - 0.0);
            const double kI = st->light[C2(st, i, j)];
            double dPO4 = 0.0;
            if (lPO4 > NULLSMALL) dPO4 = p->dtyr * ficefree * kI * kPO4 * p->k0_PO4;
            double DOMfrac = p->DOMfrac, RDOMfrac = 0.0;
            double DOMtotal = DOMfrac + RDOMfrac;
            if (DOMtotal > 1.0) {
                DOMfrac = DOMfrac / DOMtotal;
                DOMtotal = 1.0;
            }
            double red_POC_CaCO3 = 0.0;
            if (p->caco3) {
                /* 'Ridgwelletal2007ab' */
                if (ohm > 1.0)
                    red_POC_CaCO3 = (1.0 - DOMtotal) * p->red_POC_CaCO3 * pow(ohm - 1.0, p->red_POC_CaCO3_pP);
                else
                    red_POC_CaCO3 = 0.0;
            }
            const double red_POC_POP = 1.0 / p->red_POP_POC; /* biogem_data.f90:602 */
            const double kPO4f = lPO4 / (lPO4 + p->POC_c0frac2);
            const double fr2POC = (1.0 - kPO4f) * p->POC_dfrac2 + p->POC_frac2;
            const int klo = k_mld < kb ? kb : k_mld;
            for (int k = klo; k <= maxk; k++) {
                double bp[OG_NSED];
                bp[OG_S_POC] = p->red_POP_POC * dPO4;
                bp[OG_S_CACO3] = p->caco3 ? red_POC_CaCO3 * bp[OG_S_POC] : 0.0;
                bp[OG_S_POP] = red_POC_POP * bp[OG_S_POC];
                bp[OG_S_POCFR2] = 0.0;
                bp[OG_S_CACO3FR2] = 0.0;
                /* inorganic uptake, 1259-1267 (sed order POC, POP, CaCO3) */
                double upt[4] = {0, 0, 0, 0};
                for (int s = 0; s < OG_NSED; s++) {
                    if (s == OG_S_CACO3 && !p->caco3) continue;
                    for (int q = 0; q < 4; q++)
                        if (cv.c[s][q] != 0.0) upt[q] = upt[q] + cv.c[s][q] * bp[s];
                }
                /* DOM production, 1316-1357: POC -> DOM_C, POP -> DOM_P */
                double dom[OG_NSED] = {0, 0, 0, 0, 0};
                dom[OG_S_POC] = 1.0 * DOMfrac * bp[OG_S_POC];
                dom[OG_S_POP] = 1.0 * DOMfrac * bp[OG_S_POP];
                if (p->l_DOMC >= 0) {
                    double *r = &vdocn[O3(st, p->l_DOMC, i, j, k)];
                    *r = *r + dom[OG_S_POC];
                }
                if (p->l_DOMP >= 0) {
                    double *r = &vdocn[O3(st, p->l_DOMP, i, j, k)];
                    *r = *r + dom[OG_S_POP];
                }
                for (int s = 0; s < OG_NSED; s++) bp[s] = bp[s] - (dom[s] + 0.0);
                bp[OG_S_POCFR2] = fr2POC;
                bp[OG_S_CACO3FR2] = p->caco3 ? p->CaCO3_frac2 : 0.0;
                for (int s = 0; s < OG_NSED; s++) st->part[P3(st, s, i, j, k)] = bp[s];
                /* 1437-1440: uptake is negative remineralisation */
                for (int q = 0; q < 4; q++) {
                    if (cv.slot[q] < 0) continue;
                    double *r = &vdocn[O3(st, cv.slot[q], i, j, k)];
                    *r = *r - upt[q];
                }
            }
        }
    /* vdocn = bio_remin + dtyr rM locijk_focn (biogem.f90:1821-1833), the
     * surface locijk_focn being minus the ocean -> atmosphere flux */
    if (p->airsea)
        for (int i = 1; i <= maxi; i++)
            for (int j = 1; j <= maxj; j++) {
                if (KK1(i, j) > maxk) continue;
                const size_t c = C2(st, i, j);
                const double dtrM = p->dtyr * st->rM[C3(st, i, j, maxk)];
                if (p->l_DIC >= 0 && p->l_ALK >= 0) {
                    double *r = &vdocn[O3(st, p->l_DIC, i, j, maxk)];
                    *r = *r + dtrM * (0.0 - st->fatm[2 * c]);
                }
                if (p->l_O2 >= 0) {
                    double *r = &vdocn[O3(st, p->l_O2, i, j, maxk)];
                    *r = *r + dtrM * (0.0 - st->fatm[2 * c + 1]);
                }
            }
    free(vrem);
}

void og_tracercoupling(og_bgc_state *st, const og_grid *g, const int32_t *k1,
                       double *go_ts, const double *vdocn)
{
    const int maxi = g->maxi, maxj = g->maxj, maxk = g->maxk, maxl = st->maxl;
    const double off[2] = {CONST_ZEROC, SALN0};
    size_t ncell = (size_t)maxi * maxj * maxk;
    /* 1927-1941 */
    double totM = 0.0, totV = 0.0;
    for (int i = 1; i <= maxi; i++)
        for (int j = 1; j <= maxj; j++) {
            int kb = KK1(i, j);
            if (kb > maxk) continue;
            double pm = 0.0, pv = 0.0;
            for (int k = kb; k <= maxk; k++) pm += st->M[C3(st, i, j, k)];
            for (int k = kb; k <= maxk; k++) pv += st->V[C3(st, i, j, k)];
            totM += pm;
            totV += pv;
        }
    (void)totM;
    const double rV = 1.0 / totV;
    /* 1952-1959 */
    const double Sold = Smean_of(st, k1, &st->ocn[1], maxl, rV);
    const double rSold = 1.0 / Sold;
    /* 1961-1969 */
    double *OLD = calloc(maxl, sizeof(double)), *NEW = calloc(maxl, sizeof(double)), *rNEW = calloc(maxl, sizeof(double));
    double *lv = calloc(ncell * maxl, sizeof(double));
    for (int l = 2; l < maxl; l++) {
        double tot = 0.0;
        for (int i = 1; i <= maxi; i++)
            for (int j = 1; j <= maxj; j++) {
                int kb = KK1(i, j);
                if (kb > maxk) continue;
                double pt = 0.0;
                for (int k = kb; k <= maxk; k++) pt += st->ocn[O3(st, l, i, j, k)] * st->M[C3(st, i, j, k)];
                tot += pt;
            }
        OLD[l] = tot;
    }
    /* 1975-1981 */
    for (int i = 1; i <= maxi; i++)
        for (int j = 1; j <= maxj; j++)
            for (int k = KK1(i, j); k <= maxk; k++) {
                for (int l = 0; l < 2; l++) lv[O3(st, l, i, j, k)] = go_ts[O3(st, l, i, j, k)] + off[l];
                for (int l = 2; l < maxl; l++)
                    lv[O3(st, l, i, j, k)] = go_ts[O3(st, l, i, j, k)] * st->ocn[O3(st, 1, i, j, k)] * rSold;
            }
    /* 1984-1997 */
    for (int l = 2; l < maxl; l++) {
        double tot = 0.0;
        for (int i = 1; i <= maxi; i++)
            for (int j = 1; j <= maxj; j++) {
                int kb = KK1(i, j);
                if (kb > maxk) continue;
                double pt = 0.0;
                for (int k = kb; k <= maxk; k++) pt += lv[O3(st, l, i, j, k)] * st->M[C3(st, i, j, k)];
                tot += pt;
            }
        NEW[l] = tot;
        rNEW[l] = fabs(tot) < NULLSMALL ? 0.0 : 1.0 / tot;
    }
    /* 2001-2013 (ctrl_force_GOLDSTEInTS) */
    for (int i = 1; i <= maxi; i++)
        for (int j = 1; j <= maxj; j++)
            for (int k = KK1(i, j); k <= maxk; k++)
                for (int l = 0; l < 2; l++) {
                    double v = go_ts[O3(st, l, i, j, k)] + off[l] + vdocn[O3(st, l, i, j, k)];
                    lv[O3(st, l, i, j, k)] = v;
                    st->ocn[O3(st, l, i, j, k)] = v;
                }
    /* 2016-2025 */
    const double Snew = Smean_of(st, k1, &st->ocn[1], maxl, rV);
    const double Sratio = Snew / Sold, rSratio = 1.0 / Sratio;
    /* 2033-2048 */
    for (int i = 1; i <= maxi; i++)
        for (int j = 1; j <= maxj; j++)
            for (int k = KK1(i, j); k <= maxk; k++) {
                for (int l = 2; l < maxl; l++) {
                    double v = (OLD[l] * rNEW[l]) * lv[O3(st, l, i, j, k)] + vdocn[O3(st, l, i, j, k)];
                    st->ocn[O3(st, l, i, j, k)] = Sratio * v;
                }
                for (int s = 0; s < OG_NSED; s++) {
                    double *pp = &st->part[P3(st, s, i, j, k)];
                    *pp = *pp + 0.0;
                    *pp = Sratio * *pp;
                }
                for (int l = 0; l < 2; l++) go_ts[O3(st, l, i, j, k)] = lv[O3(st, l, i, j, k)] - off[l];
                for (int l = 2; l < maxl; l++)
                    go_ts[O3(st, l, i, j, k)] = (Snew / st->ocn[O3(st, 1, i, j, k)]) * st->ocn[O3(st, l, i, j, k)];
            }
    /* 2049-2061 */
    for (int i = 1; i <= maxi; i++)
        for (int j = 1; j <= maxj; j++)
            for (int k = KK1(i, j); k <= maxk; k++) {
                st->M[C3(st, i, j, k)] = rSratio * st->M[C3(st, i, j, k)];
                st->rM[C3(st, i, j, k)] = Sratio * st->rM[C3(st, i, j, k)];
            }
    free(OLD); free(NEW); free(rNEW); free(lv);
}

/* ------------------------------------------------------------------------ */
/* GOLDSTEIN momentum: the velocity solve                                    */
/* ------------------------------------------------------------------------ */
static const double MOM_USC = 0.05, MOM_FSC = 2 * 7.2921E-5, MOM_RH0SC = 1.0E3; /* goldstein_lib.f90:49-50 */

#define MI (mo->maxi)
#define MJ (mo->maxj)
#define MK1(i, j) (k1[(i) + (MI + 2) * (j)])
#define MH(l, i, j) (mo->h[(size_t)((l) - 1) + 3 * ((size_t)(i) + (size_t)(MI + 2) * (size_t)(j))])
#define MRH(l, i, j) (mo->rh[(size_t)((l) - 1) + 3 * ((size_t)(i) + (size_t)(MI + 2) * (size_t)(j))])
#define MKU(l, i, j) (mo->ku[(size_t)((l) - 1) + 2 * ((size_t)((i) - 1) + (size_t)MI * (size_t)((j) - 1))])
#define MMK(i, j) (mo->mk[(size_t)(i) + (size_t)(MI + 2) * (size_t)(j)])
#define MGETJ(i, j) (mo->getj[(size_t)((i) - 1) + (size_t)MI * (size_t)((j) - 1)])
#define MDRAG(l, i, j) (mo->drag[(size_t)((l) - 1) + 2 * ((size_t)(i) + (size_t)(MI + 2) * (size_t)(j))])
#define MRTV(i, j) (mo->rtv[(size_t)((i) - 1) + (size_t)MI * (size_t)((j) - 1)])
#define MRTV3(i, j) (mo->rtv3[(size_t)((i) - 1) + (size_t)MI * (size_t)((j) - 1)])
#define MGAP(k, l) (mo->gap[(size_t)((k) - 1) * (size_t)(2 * MI + 3) + (size_t)((l) - 1)])
#define MRATM(k, l) (mo->ratm[(size_t)((k) - 1) * (size_t)(MI + 1) + (size_t)((l) - 1)])
#define MGB(k) (mo->gb[(k) - 1])
#define MGBOLD(k) (mo->gbold[(k) - 1])
#define MBP(i, j, k) (mo->bp[(size_t)(i) + (size_t)(MI + 2) * ((size_t)(j) + (size_t)(MJ + 2) * (size_t)(k))])
#define MSBP(i, j) (mo->sbp[(size_t)(i) + (size_t)(MI + 2) * (size_t)(j)])
#define UBL(ubv, l, i, j) ((ubv)[(size_t)((l) - 1) + 2 * ((size_t)(i) + (size_t)(MI + 2) * (size_t)(j))])
#define PSIL(psv, i, j) ((psv)[(size_t)(i) + (size_t)(MI + 1) * (size_t)(j)])
#define MERISL(a, b) (mo->erisl[(size_t)((a) - 1) + (size_t)mo->isles * (size_t)((b) - 1)])
#define MTAU(l, i, j) (mo->tau[(size_t)((l) - 1) + 2 * ((size_t)((i) - 1) + (size_t)MI * (size_t)((j) - 1))])
#define MDZTAU(l, i, j) (mo->dztau[(size_t)((l) - 1) + 2 * ((size_t)((i) - 1) + (size_t)MI * (size_t)((j) - 1))])
#define MDZTAV(l, i, j) (mo->dztav[(size_t)((l) - 1) + 2 * ((size_t)((i) - 1) + (size_t)MI * (size_t)((j) - 1))])
#define MU(uv, l, i, j, k) ((uv)[(size_t)((l) - 1) + 3 * ((size_t)(i) + (size_t)(MI + 1) * ((size_t)(j) + (size_t)(MJ + 1) * (size_t)((k) - 1)))])
#define MRHO(i, j, k) (rho[(size_t)(i) + (size_t)(MI + 2) * ((size_t)(j) + (size_t)(MJ + 2) * (size_t)(k))])
#define ISIGN1(x) ((x) >= 0 ? 1 : -1)

static int imax2(int a, int b) { return a > b ? a : b; }
static int imin2(int a, int b) { return a < b ? a : b; }

/* drgset, goldstein.f90:2845-2885: drag at psi points, raised twice by drgf
 * near the equator and over shallow water (k1 > kmxdrg), then interpolated
 * to the velocity points */
static void mom_drgset(og_mom *mo, const int32_t *k1, double adrag, double drgf, int kmxdrg, int jeb)
{
    double *tmp = calloc((size_t)(MI + 1) * (MJ + 1), sizeof(double));
#define TMPD(i, j) tmp[(size_t)(i) + (size_t)(MI + 1) * (size_t)(j)]
    for (int j = 0; j <= MJ; j++)
        for (int i = 0; i <= MI; i++) {
            const int kloc2 = imax2(imax2(MK1(i, j), MK1(i + 1, j)), imax2(MK1(i, j + 1), MK1(i + 1, j + 1)));
            int kloc4 = MK1(i, j);
            for (int j1 = imax2(0, j - 1); j1 <= imin2(MJ + 1, j + 2); j1++)
                for (int i1 = i - 1; i1 <= i + 2; i1++) {
                    const int i1p = 1 + (MI + i1 - 1) % MI;
                    kloc4 = imax2(kloc4, MK1(i1p, j1));
                }
            if (kloc2 > kmxdrg || abs(j - MJ / 2) <= jeb)
                TMPD(i, j) = adrag * drgf * drgf;
            else if (kloc4 > kmxdrg || abs(j - MJ / 2) == jeb + 1)
                TMPD(i, j) = adrag * drgf;
            else
                TMPD(i, j) = adrag;
        }
    for (int j = 1; j <= MJ; j++)
        for (int i = 1; i <= MI; i++) {
            MDRAG(1, i, j) = 0.5 * (TMPD(i, j) + TMPD(i, j - 1));
            MDRAG(2, i, j) = 0.5 * (TMPD(i, j) + TMPD(i - 1, j));
        }
    for (int j = 0; j <= MJ + 1; j++) MDRAG(2, MI + 1, j) = MDRAG(2, 1, j);
#undef TMPD
    free(tmp);
}

/* invert, goldstein.f90:3138-3204: the psi equation at psi points (i, j),
 * i = 1..maxi (periodic), j = 0..maxj, as a band of half-width maxi+1, and
 * its elimination without pivoting (multipliers kept in ratm) */
static void mom_invert(og_mom *mo, const og_grid *g, const int32_t *k1)
{
    const int n = MI, m = MJ + 1, W = 2 * n + 3;
    memset(mo->gap, 0, sizeof(double) * (size_t)n * m * W);
    for (int i = 1; i <= MI; i++)
        for (int j = 0; j <= MJ; j++) {
            const int k = i + j * n;
            if (imax2(imax2(MK1(i, j), MK1(i + 1, j)), imax2(MK1(i, j + 1), MK1(i + 1, j + 1))) <= g->maxk) {
                const double tv = (g->s[j + 1] * MRH(1, i, j + 1) - g->s[j] * MRH(1, i, j)) / (2.0 * g->dsv[j] * g->dphi);
                const double tv1 = (g->sv[j] * MRH(2, i + 1, j) - g->sv[j] * MRH(2, i, j)) / (2.0 * g->dphi * g->dsv[j]);
                MGAP(k, 2) = MDRAG(1, i, j) * g->c[j] * g->c[j] * MRH(1, i, j) / (g->ds[j] * g->dsv[j]) + tv1;
                int l = n + 1;
                if (i == 1) l = 2 * n + 1;
                MGAP(k, l) = MDRAG(2, i, j) * g->rcv[j] * g->rcv[j] * g->rdphi * g->rdphi * MRH(2, i, j) - tv;
                MGAP(k, n + 2) = -(MDRAG(2, i, j) * MRH(2, i, j) + MDRAG(2, i + 1, j) * MRH(2, i + 1, j)) /
                                     (g->cv[j] * g->cv[j] * g->dphi * g->dphi) -
                                 (MDRAG(1, i, j) * g->c[j] * g->c[j] * MRH(1, i, j) / g->ds[j] +
                                  MDRAG(1, i, j + 1) * g->c[j + 1] * g->c[j + 1] * MRH(1, i, j + 1) / g->ds[j + 1]) /
                                     g->dsv[j];
                l = n + 3;
                if (i == MI) l = 3;
                MGAP(k, l) = MDRAG(2, i + 1, j) * MRH(2, i + 1, j) / (g->cv[j] * g->cv[j] * g->dphi * g->dphi) + tv;
                MGAP(k, 2 * n + 2) = MDRAG(1, i, j + 1) * g->c[j + 1] * g->c[j + 1] * MRH(1, i, j + 1) /
                                         (g->ds[j + 1] * g->dsv[j]) -
                                     tv1;
            } else {
                MGAP(k, n + 2) = 1;
            }
        }
    memset(mo->ratm, 0, sizeof(double) * (size_t)n * m * (n + 1));
    for (int i = 1; i <= n * m - 1; i++) {
        const int im = imin2(i + n + 1, n * m);
        for (int j = i + 1; j <= im; j++) {
            const double rat = MGAP(j, n + 2 - j + i) / MGAP(i, n + 2);
            MRATM(j, j - i) = rat;
            if (rat != 0)
                for (int k = n + 2 - j + i; k <= 2 * n + 3 - j + i; k++)
                    MGAP(j, k) = MGAP(j, k) - rat * MGAP(i, k + j - i);
        }
    }
}

/* ubarsolv, goldstein.f90:3500-3565: forward elimination and back
 * substitution of gb, psi from it, the barotropic velocities and their
 * boundary values */
static void mom_ubarsolv(og_mom *mo, const og_grid *g, double *ubv, double *psv)
{
    const int n = MI, m = MJ + 1;
    for (int i = 1; i <= n * m - 1; i++) {
        const int im = imin2(i + n + 1, n * m);
        for (int j = i + 1; j <= im; j++) MGB(j) = MGB(j) - MRATM(j, j - i) * MGB(i);
    }
    MGB(n * m) = MGB(n * m) / MGAP(n * m, n + 2);
    for (int i = n * m - 1; i >= 1; i--) {
        const int km = imin2(n + 1, n * m - i);
        for (int k = 1; k <= km; k++) MGB(i) = MGB(i) - MGAP(i, n + 2 + k) * MGB(i + k);
        MGB(i) = MGB(i) / MGAP(i, n + 2);
    }
    for (int j = 0; j <= MJ; j++) {
        for (int i = 1; i <= MI; i++) PSIL(psv, i, j) = MGB(i + j * n);
        PSIL(psv, 0, j) = PSIL(psv, MI, j);
    }
    for (int j = 1; j <= MJ; j++)
        for (int i = 1; i <= MI; i++)
            UBL(ubv, 1, i, j) = -MRH(1, i, j) * g->c[j] * (PSIL(psv, i, j) - PSIL(psv, i, j - 1)) * g->rds[j];
    for (int j = 1; j <= MJ - 1; j++)
        for (int i = 1; i <= MI; i++)
            UBL(ubv, 2, i, j) = MRH(2, i, j) * (PSIL(psv, i, j) - PSIL(psv, i - 1, j)) * g->rcv[j] * g->rdphi;
    for (int i = 1; i <= MI; i++) {
        UBL(ubv, 2, i, MJ) = 0.0;
        UBL(ubv, 2, i, 0) = 0.0;
    }
    for (int j = 1; j <= MJ; j++) {
        UBL(ubv, 2, MI + 1, j) = UBL(ubv, 2, 1, j);
        UBL(ubv, 1, 0, j) = UBL(ubv, 1, MI, j);
        UBL(ubv, 1, MI + 1, j) = UBL(ubv, 1, 1, j);
        UBL(ubv, 2, 0, j) = UBL(ubv, 2, MI, j);
    }
    UBL(ubv, 2, MI + 1, 0) = UBL(ubv, 2, 1, 0);
    UBL(ubv, 2, 0, 0) = UBL(ubv, 2, MI, 0);
}

/* island, goldstein_lib.f90:186-241: the path integral round island isl of
 * the barotropic momentum balance (indj = 1 adds wind and the bottom-pressure
 * torque) */
static double mom_island(const og_mom *mo, const og_grid *g, const double *ubv, int isl, int indj)
{
    double erisl1 = 0.0;
    for (int p = 1; p <= mo->npi[isl - 1]; p++) {
        const size_t q = (size_t)(p - 1) + (size_t)mo->mpi * (size_t)(isl - 1);
        const int lpi = mo->lpisl[q], ipi = mo->ipisl[q], jpi = mo->jpisl[q];
        const int al = abs(lpi);
        double cor;
        if (al == 1)
            cor = -g->s[jpi] * 0.25 *
                  (UBL(ubv, 2, ipi, jpi) + UBL(ubv, 2, ipi + 1, jpi) + UBL(ubv, 2, ipi, jpi - 1) + UBL(ubv, 2, ipi + 1, jpi - 1));
        else
            cor = g->sv[jpi] * 0.25 *
                  (UBL(ubv, 1, ipi - 1, jpi) + UBL(ubv, 1, ipi, jpi) + UBL(ubv, 1, ipi - 1, jpi + 1) + UBL(ubv, 1, ipi, jpi + 1));
        const double sgn = (double)ISIGN1(lpi);
        const double rhv = al == 1 ? MRH(1, ipi, jpi) : MRH(2, ipi, jpi);
        const double tauv = indj ? (al == 1 ? MTAU(1, ipi, jpi) : MTAU(2, ipi, jpi)) : 0.0;
        erisl1 = erisl1 + sgn * (MDRAG(al, ipi, jpi) * UBL(ubv, al, ipi, jpi) + cor - (double)indj * tauv * rhv) *
                              (g->c[jpi] * g->dphi * (2.0 - al) + g->rcv[jpi] * g->dsv[jpi] * (al - 1.0));
        if (indj == 1) {
            if (al == 1) {
                double tv1 = 0.0;
                for (int k = MKU(1, ipi, jpi); k <= MMK(ipi + 1, jpi); k++) tv1 = tv1 + MBP(ipi + 1, jpi, k) * g->dz[k];
                for (int k = MKU(1, ipi, jpi); k <= MMK(ipi, jpi); k++) tv1 = tv1 - MBP(ipi, jpi, k) * g->dz[k];
                erisl1 = erisl1 + (MSBP(ipi + 1, jpi) - MSBP(ipi, jpi) + tv1) * sgn * MRH(1, ipi, jpi);
            } else {
                double tv2 = 0.0;
                for (int k = MKU(2, ipi, jpi); k <= MMK(ipi, jpi + 1); k++) tv2 = tv2 + MBP(ipi, jpi + 1, k) * g->dz[k];
                for (int k = MKU(2, ipi, jpi); k <= MMK(ipi, jpi); k++) tv2 = tv2 - MBP(ipi, jpi, k) * g->dz[k];
                erisl1 = erisl1 + (MSBP(ipi, jpi + 1) - MSBP(ipi, jpi) + tv2) * sgn * MRH(2, ipi, jpi);
            }
        }
    }
    return erisl1;
}

/* matinv_gold and matmult, goldstein.f90:3452-3492 (column-major amat) */
static void mom_matinv(int nvar, double *a)
{
#define AM(i, j) a[(size_t)((i) - 1) + (size_t)nvar * (size_t)((j) - 1)]
    for (int i = 1; i <= nvar - 1; i++)
        for (int j = i + 1; j <= nvar; j++)
            for (int k = i + 1; k <= nvar; k++) AM(j, k) = AM(i, i) * AM(j, k) - AM(j, i) * AM(i, k);
}
static void mom_matmult(int nvar, const double *a, double *rhs)
{
    for (int i = 1; i <= nvar - 1; i++)
        for (int j = i + 1; j <= nvar; j++) rhs[j - 1] = AM(i, i) * rhs[j - 1] - AM(j, i) * rhs[i - 1];
    rhs[nvar - 1] = rhs[nvar - 1] / AM(nvar, nvar);
    for (int i = nvar - 1; i >= 1; i--) {
        for (int j = i + 1; j <= nvar; j++) rhs[i - 1] = rhs[i - 1] - AM(i, j) * rhs[j - 1];
        rhs[i - 1] = rhs[i - 1] / AM(i, i);
    }
#undef AM
}

void og_mom_free(og_mom *mo)
{
    free(mo->h); free(mo->rh); free(mo->ku); free(mo->mk); free(mo->getj); free(mo->drag);
    free(mo->rtv); free(mo->rtv3); free(mo->gap); free(mo->ratm); free(mo->gb); free(mo->gbold);
    free(mo->bp); free(mo->sbp); free(mo->ub); free(mo->psi); free(mo->ubisl); free(mo->psisl);
    free(mo->erisl); free(mo->psibc); free(mo->npi); free(mo->lpisl); free(mo->ipisl); free(mo->jpisl);
    free(mo->tau); free(mo->dztau); free(mo->dztav); free(mo->u1);
    memset(mo, 0, sizeof(*mo));
}

int og_mom_init(og_mom *mo, const og_grid *g, const int32_t *k1, const int32_t *labels, int isles,
                const int32_t *npi, const int32_t *paths, double adrag, double rel)
{
    memset(mo, 0, sizeof(*mo));
    mo->maxi = g->maxi;
    mo->maxj = g->maxj;
    mo->maxk = g->maxk;
    mo->isles = isles;
    mo->rel = rel;
    const int maxk = g->maxk, n = MI, N = MI * (MJ + 1);
    const size_t P2 = (size_t)(MI + 2) * (MJ + 2);
    mo->h = calloc(3 * P2, sizeof(double));
    mo->rh = calloc(3 * P2, sizeof(double));
    mo->ku = calloc((size_t)2 * MI * MJ, sizeof(int32_t));
    mo->mk = calloc(P2, sizeof(int32_t));
    mo->getj = calloc((size_t)MI * MJ, 1);
    mo->drag = calloc(2 * P2, sizeof(double));
    mo->rtv = calloc((size_t)MI * MJ, sizeof(double));
    mo->rtv3 = calloc((size_t)MI * MJ, sizeof(double));
    mo->gap = calloc((size_t)N * (2 * n + 3), sizeof(double));
    mo->ratm = calloc((size_t)N * (n + 1), sizeof(double));
    mo->gb = calloc(N, sizeof(double));
    mo->gbold = calloc(N, sizeof(double));
    mo->bp = calloc(P2 * (maxk + 1), sizeof(double));
    mo->sbp = calloc(P2, sizeof(double));
    mo->ub = calloc((size_t)2 * (MI + 2) * (MJ + 1), sizeof(double));
    mo->psi = calloc((size_t)(MI + 1) * (MJ + 1), sizeof(double));
    mo->tau = calloc((size_t)2 * MI * MJ, sizeof(double));
    mo->dztau = calloc((size_t)2 * MI * MJ, sizeof(double));
    mo->dztav = calloc((size_t)2 * MI * MJ, sizeof(double));
    mo->u1 = calloc((size_t)3 * (MI + 1) * (MJ + 1) * maxk, sizeof(double));
    /* seabed depth and its reciprocal at T, u and v points (1359-1384) */
    for (int j = MJ + 1; j >= 0; j--)
        for (int i = 0; i <= MI + 1; i++)
            if (MK1(i, j) <= maxk) {
                for (int k = MK1(i, j); k <= maxk; k++) MH(3, i, j) = MH(3, i, j) + g->dz[k];
                MRH(3, i, j) = 1.0 / MH(3, i, j);
            }
    for (int j = 0; j <= MJ + 1; j++)
        for (int i = 0; i <= MI; i++) {
            MH(1, i, j) = fmin(MH(3, i, j), MH(3, i + 1, j));
            if (imax2(MK1(i, j), MK1(i + 1, j)) <= maxk) MRH(1, i, j) = 1.0 / MH(1, i, j);
        }
    for (int j = 0; j <= MJ; j++)
        for (int i = 0; i <= MI + 1; i++) {
            MH(2, i, j) = fmin(MH(3, i, j), MH(3, i, j + 1));
            if (imax2(MK1(i, j), MK1(i, j + 1)) <= maxk) MRH(2, i, j) = 1.0 / MH(2, i, j);
        }
    for (int j = 1; j <= MJ; j++)
        for (int i = 1; i <= MI; i++) {
            MKU(1, i, j) = imax2(MK1(i, j), MK1(i + 1, j));
            MKU(2, i, j) = imax2(MK1(i, j), MK1(i, j + 1));
        }
    /* drag (1395-1405) and the reciprocal denominators (1409-1414) */
    mom_drgset(mo, k1, 1.0 / (adrag * 86400 * MOM_FSC), 3.0, maxk / 2, 1);
    for (int j = 1; j <= MJ; j++)
        for (int i = 1; i <= MI; i++) {
            MRTV(i, j) = 1.0 / (g->s[j] * g->s[j] + MDRAG(1, i, j) * MDRAG(1, i, j));
            MRTV3(i, j) = 1.0 / (g->sv[j] * g->sv[j] + MDRAG(2, i, j) * MDRAG(2, i, j));
        }
    /* mk and getj (1460-1493) */
#define WETF(i, j) ((1 + ISIGN1(maxk - MK1(i, j))) / 2)
    for (int j = 1; j <= MJ; j++)
        for (int i = 1; i <= MI; i++) {
            int v = imax2(imax2(MK1(i, j) * WETF(i, j), MK1(i + 1, j) * WETF(i + 1, j)),
                          imax2(imax2(MK1(i - 1, j) * WETF(i - 1, j), MK1(i, j + 1) * WETF(i, j + 1)),
                                MK1(i, j - 1) * WETF(i, j - 1)));
            MMK(i, j) = v * WETF(i, j);
        }
#undef WETF
    for (int j = 1; j <= MJ; j++) MMK(MI + 1, j) = MMK(1, j);
    for (int j = 1; j <= MJ; j++)
        for (int i = 1; i <= MI; i++)
            MGETJ(i, j) = (imax2(imax2(MK1(i, j), MK1(i + 1, j)), imax2(MK1(i, j + 1), MK1(i + 1, j + 1))) <= maxk) &&
                          (MK1(i, j) != MK1(i, j + 1) || MK1(i, j) != MK1(i + 1, j) || MK1(i, j) != MK1(i + 1, j + 1));
    /* islands (1505-1575): labels as gbold, the paths checked as the reference checks them */
    for (int j = 0; j <= MJ; j++)
        for (int i = 1; i <= MI; i++) MGBOLD(i + j * MI) = (double)labels[(i - 1) + (size_t)MI * j];
    int mpi = 1;
    for (int q = 0; q < isles; q++) mpi = imax2(mpi, npi[q]);
    mo->mpi = mpi;
    const int nis = isles > 0 ? isles : 1;
    mo->npi = calloc(nis, sizeof(int32_t));
    mo->lpisl = calloc((size_t)mpi * nis, sizeof(int32_t));
    mo->ipisl = calloc((size_t)mpi * nis, sizeof(int32_t));
    mo->jpisl = calloc((size_t)mpi * nis, sizeof(int32_t));
    mo->psisl = calloc((size_t)(MI + 1) * (MJ + 1) * nis, sizeof(double));
    mo->ubisl = calloc((size_t)2 * (MI + 2) * (MJ + 1) * nis, sizeof(double));
    mo->erisl = calloc((size_t)nis * (nis + 1), sizeof(double));
    mo->psibc = calloc(nis, sizeof(double));
    size_t o = 0;
    for (int q = 0; q < isles; q++) {
        mo->npi[q] = npi[q];
        for (int p = 0; p < npi[q]; p++, o++) {
            const int lpi = paths[3 * o], ipi = paths[3 * o + 1], jpi = paths[3 * o + 2];
            /* the reference's checks (1555-1570), and the index ranges its
             * arrays give island(): ku(2, maxi, maxj), ub(2, 0:maxi+1, 0:maxj) */
            if ((abs(lpi) != 1 && abs(lpi) != 2) || ipi > MI || ipi < 1 || jpi > MJ || jpi < 1 || MK1(ipi, jpi) > maxk ||
                (abs(lpi) == 2 && jpi + 1 > MJ))
                return -1;
            mo->lpisl[p + (size_t)mpi * q] = lpi;
            mo->ipisl[p + (size_t)mpi * q] = ipi;
            mo->jpisl[p + (size_t)mpi * q] = jpi;
        }
    }
    /* the factorisation and the island basis (1788-1814) */
    mom_invert(mo, g, k1);
    for (int isol = 1; isol <= isles; isol++) {
        for (int j = 0; j <= MJ; j++)
            for (int i = 1; i <= MI; i++) {
                const int k = i + j * MI;
                MGB(k) = ((int)MGBOLD(k) == isol + 1) ? 1.0 : 0.0;
            }
        double *ubv = mo->ubisl + (size_t)2 * (MI + 2) * (MJ + 1) * (isol - 1);
        double *psv = mo->psisl + (size_t)(MI + 1) * (MJ + 1) * (isol - 1);
        mom_ubarsolv(mo, g, ubv, psv);
        for (int isl = 1; isl <= isles; isl++) MERISL(isl, isol) = mom_island(mo, g, ubv, isl, 0);
    }
    if (isles > 0) mom_matinv(isles, mo->erisl);
    return 0;
}

void og_mom_set_wind(og_mom *mo, const og_grid *g, const double *stress, double scf)
{
    const double dzz = g->dz[g->maxk] * g->dza[g->maxk - 1] / 2;
    const double d = MOM_RH0SC * DSC * MOM_USC * MOM_FSC;
    for (int j = 1; j <= MJ; j++)
        for (int i = 1; i <= MI; i++) {
            const size_t c = (size_t)4 * ((i - 1) + (size_t)MI * (j - 1));
            MDZTAU(1, i, j) = scf * stress[c + 0] / d / dzz;
            MDZTAU(2, i, j) = scf * stress[c + 1] / d / dzz;
            MDZTAV(1, i, j) = scf * stress[c + 2] / d / dzz;
            MDZTAV(2, i, j) = scf * stress[c + 3] / d / dzz;
            MTAU(1, i, j) = MDZTAU(1, i, j) * dzz;
            MTAU(2, i, j) = MDZTAV(2, i, j) * dzz;
        }
}

void og_rho_from_ts(const og_grid *g, int maxl, const int32_t *k1, const double *go_ts, double *rho)
{
    og_mom tmp;
    og_mom *mo = &tmp;
    mo->maxi = g->maxi;
    mo->maxj = g->maxj;
    for (int k = 1; k <= g->maxk; k++)
        for (int j = 1; j <= MJ; j++) {
            for (int i = 1; i <= MI; i++)
                if (k >= MK1(i, j)) {
                    const size_t c = (size_t)maxl * ((i - 1) + (size_t)MI * ((j - 1) + (size_t)MJ * (k - 1)));
                    MRHO(i, j, k) = og_eos(g->ec, go_ts[c], go_ts[c + 1]);
                }
            MRHO(0, j, k) = MRHO(MI, j, k);
            MRHO(MI + 1, j, k) = MRHO(1, j, k);
        }
}

void og_mom_step(og_mom *mo, const og_grid *g, const int32_t *k1, const double *rho, double *u)
{
    const int maxk = g->maxk, n = MI;
    /* wind, 3685-3714 */
    for (int i = 1; i <= MI; i++)
        for (int j = 0; j <= MJ; j++) {
            const int ip1 = i % MI + 1;
            const int k = i + j * n;
            if (imax2(imax2(MK1(i, j), MK1(i + 1, j)), imax2(MK1(i, j + 1), MK1(i + 1, j + 1))) <= maxk) {
                MGB(k) = (MTAU(2, ip1, j) * MRH(2, i + 1, j) - MTAU(2, i, j) * MRH(2, i, j)) * g->rdphi * g->rcv[j] -
                         (MTAU(1, i, j + 1) * g->c[j + 1] * MRH(1, i, j + 1) - MTAU(1, i, j) * g->c[j] * MRH(1, i, j)) *
                             g->rdsv[j];
            } else {
                MGB(k) = 0;
            }
            MGBOLD(k) = MGB(k);
        }
    /* jbar, 3217-3315 */
    for (int j = 1; j <= MJ; j++)
        for (int i = 1; i <= MI; i++)
            if (MK1(i, j) <= maxk)
                for (int k = MK1(i, j) + 1; k <= maxk; k++)
                    MBP(i, j, k) = MBP(i, j, k - 1) - (MRHO(i, j, k) + MRHO(i, j, k - 1)) * g->dza[k - 1] * 0.5;
    for (int j = 1; j <= MJ; j++)
        for (int i = 1; i <= MI; i++)
            if (MMK(i, j) > 0) {
                MSBP(i, j) = 0;
                for (int k = MMK(i, j) + 1; k <= maxk; k++) MSBP(i, j) = MSBP(i, j) + MBP(i, j, k) * g->dz[k];
            }
    for (int j = 1; j <= MJ; j++) {
        if (MK1(1, j) < maxk)
            for (int k = MK1(1, j); k <= maxk; k++) MBP(MI + 1, j, k) = MBP(1, j, k);
        if (MK1(1, j) <= maxk) MSBP(MI + 1, j) = MSBP(1, j);
    }
    for (int j = 1; j <= MJ - 1; j++)
        for (int i = 1; i <= MI; i++) {
            const int ip1 = i % MI + 1;
            const int l = i + j * n;
            if (MGETJ(i, j)) {
                double tv1 = 0, tv2 = 0, tv3 = 0, tv4 = 0;
                for (int k = MKU(2, ip1, j); k <= MMK(ip1, j + 1); k++) tv1 = tv1 + MBP(ip1, j + 1, k) * g->dz[k];
                for (int k = MKU(2, ip1, j); k <= MMK(ip1, j); k++) tv2 = tv2 + MBP(ip1, j, k) * g->dz[k];
                for (int k = MKU(2, i, j); k <= MMK(i, j + 1); k++) tv3 = tv3 + MBP(i, j + 1, k) * g->dz[k];
                for (int k = MKU(2, i, j); k <= MMK(i, j); k++) tv4 = tv4 + MBP(i, j, k) * g->dz[k];
                MGB(l) = MGBOLD(l) + ((tv3 + MSBP(i, j + 1) - tv4 - MSBP(i, j)) * MRH(2, i, j) -
                                      (tv1 + MSBP(ip1, j + 1) - tv2 - MSBP(ip1, j)) * MRH(2, ip1, j)) *
                                         g->rdphi * g->rdsv[j];
                tv1 = 0;
                for (int k = MKU(1, i, j + 1); k <= MMK(ip1, j + 1); k++) tv1 = tv1 + MBP(ip1, j + 1, k) * g->dz[k];
                tv2 = 0;
                for (int k = MKU(1, i, j); k <= MMK(ip1, j); k++) tv2 = tv2 + MBP(ip1, j, k) * g->dz[k];
                tv3 = 0;
                for (int k = MKU(1, i, j + 1); k <= MMK(i, j + 1); k++) tv3 = tv3 + MBP(i, j + 1, k) * g->dz[k];
                tv4 = 0;
                for (int k = MKU(1, i, j); k <= MMK(i, j); k++) tv4 = tv4 + MBP(i, j, k) * g->dz[k];
                MGB(l) = MGB(l) + ((tv1 + MSBP(ip1, j + 1) - tv3 - MSBP(i, j + 1)) * MRH(1, i, j + 1) -
                                   (tv2 + MSBP(ip1, j) - tv4 - MSBP(i, j)) * MRH(1, i, j)) *
                                      g->rdphi * g->rdsv[j];
            } else {
                MGB(l) = MGBOLD(l);
            }
        }
    /* ubarsolv, the island correction (198-229) */
    mom_ubarsolv(mo, g, mo->ub, mo->psi);
    const int isles = mo->isles;
    if (isles > 0) {
        for (int isl = 1; isl <= isles; isl++) MERISL(isl, isles + 1) = mom_island(mo, g, mo->ub, isl, 1);
        if (isles > 1) {
            mom_matmult(isles, mo->erisl, &MERISL(1, isles + 1));
            for (int isl = 1; isl <= isles; isl++) mo->psibc[isl - 1] = -MERISL(isl, isles + 1);
        } else {
            mo->psibc[0] = -MERISL(1, 2) / MERISL(1, 1);
        }
        for (int j = 1; j <= MJ; j++)
            for (int i = 0; i <= MI + 1; i++)
                for (int l = 1; l <= 2; l++) {
                    double sum = 0.0;
                    for (int isl = 1; isl <= isles; isl++)
                        sum = sum + UBL(mo->ubisl + (size_t)2 * (MI + 2) * (MJ + 1) * (isl - 1), l, i, j) * mo->psibc[isl - 1];
                    UBL(mo->ub, l, i, j) = UBL(mo->ub, l, i, j) + sum;
                }
        for (int j = 0; j <= MJ; j++)
            for (int i = 0; i <= MI; i++) {
                double sum = 0.0;
                for (int isl = 1; isl <= isles; isl++)
                    sum = sum + PSIL(mo->psisl + (size_t)(MI + 1) * (MJ + 1) * (isl - 1), i, j) * mo->psibc[isl - 1];
                PSIL(mo->psi, i, j) = PSIL(mo->psi, i, j) + sum;
            }
    }
    /* velc, 3568-3679 */
    double dzu[2][OG_MAXK + 2];
    for (int j = 1; j <= MJ; j++)
        for (int i = 1; i <= MI; i++) {
            double sum[2] = {0, 0};
            for (int k = MK1(i, j); k <= maxk; k++) {
                double tv1, tv2, tv4, tv5;
                if (MK1(i + 1, j) > k) {
                    tv1 = 0;
                    tv2 = 0;
                } else {
                    tv2 = -(MRHO(i + 1, j, k) - MRHO(i, j, k)) * g->rdphi * g->rc[j];
                    if (imax2(imax2(MK1(i, j - 1), MK1(i, j + 1)), imax2(MK1(i + 1, j - 1), MK1(i + 1, j + 1))) <= k)
                        tv1 = -g->c[j] * (MRHO(i + 1, j + 1, k) - MRHO(i + 1, j - 1, k) + MRHO(i, j + 1, k) - MRHO(i, j - 1, k)) *
                              g->rds2[j] * 0.25;
                    else if (imax2(MK1(i, j - 1), MK1(i + 1, j - 1)) <= k)
                        tv1 = -g->c[j] * (MRHO(i + 1, j, k) - MRHO(i + 1, j - 1, k) + MRHO(i, j, k) - MRHO(i, j - 1, k)) *
                              g->rdsv[j - 1] * 0.5;
                    else if (imax2(MK1(i, j + 1), MK1(i + 1, j + 1)) <= k)
                        tv1 = -g->c[j] * (MRHO(i + 1, j + 1, k) - MRHO(i + 1, j, k) + MRHO(i, j + 1, k) - MRHO(i, j, k)) *
                              g->rdsv[j] * 0.5;
                    else
                        tv1 = 0;
                }
                if (MK1(i, j + 1) > k) {
                    tv4 = 0;
                    tv5 = 0;
                } else {
                    tv4 = -g->cv[j] * (MRHO(i, j + 1, k) - MRHO(i, j, k)) * g->rdsv[j];
                    if (imax2(imax2(MK1(i - 1, j), MK1(i - 1, j + 1)), imax2(MK1(i + 1, j), MK1(i + 1, j + 1))) <= k)
                        tv5 = -(MRHO(i + 1, j + 1, k) - MRHO(i - 1, j + 1, k) + MRHO(i + 1, j, k) - MRHO(i - 1, j, k)) *
                              g->rdphi * 0.25 * g->rcv[j];
                    else if (imax2(MK1(i - 1, j), MK1(i - 1, j + 1)) <= k)
                        tv5 = -(MRHO(i, j + 1, k) - MRHO(i - 1, j + 1, k) + MRHO(i, j, k) - MRHO(i - 1, j, k)) * g->rdphi *
                              0.5 * g->rcv[j];
                    else if (imax2(MK1(i + 1, j), MK1(i + 1, j + 1)) <= k)
                        tv5 = -(MRHO(i + 1, j + 1, k) - MRHO(i, j + 1, k) + MRHO(i + 1, j, k) - MRHO(i, j, k)) * g->rdphi *
                              0.5 * g->rcv[j];
                    else
                        tv5 = 0;
                }
                if (k == maxk) {
                    if (MK1(i + 1, j) <= k) {
                        tv1 = tv1 - MDZTAU(2, i, j);
                        tv2 = tv2 - MDZTAU(1, i, j);
                    }
                    if (MK1(i, j + 1) <= k) {
                        tv4 = tv4 - MDZTAV(2, i, j);
                        tv5 = tv5 - MDZTAV(1, i, j);
                    }
                }
                dzu[0][k] = -(g->s[j] * tv1 + MDRAG(1, i, j) * tv2) * MRTV(i, j);
                dzu[1][k] = -(MDRAG(2, i, j) * tv4 - g->sv[j] * tv5) * MRTV3(i, j);
                for (int l = 1; l <= 2; l++) {
                    if (k == MK1(i, j)) {
                        MU(u, l, i, j, k) = 0;
                    } else {
                        MU(u, l, i, j, k) = MU(u, l, i, j, k - 1) + g->dza[k - 1] * (dzu[l - 1][k] + dzu[l - 1][k - 1]) * 0.5;
                        sum[l - 1] = sum[l - 1] + g->dz[k] * MU(u, l, i, j, k);
                    }
                }
            }
            for (int k = MK1(i, j); k <= maxk; k++) {
                if (MK1(i + 1, j) <= k) {
                    MU(u, 1, i, j, k) = MU(u, 1, i, j, k) - sum[0] * MRH(1, i, j) + UBL(mo->ub, 1, i, j);
                    MU(u, 1, i, j, k) = mo->rel * MU(mo->u1, 1, i, j, k) + (1.0 - mo->rel) * MU(u, 1, i, j, k);
                    MU(mo->u1, 1, i, j, k) = MU(u, 1, i, j, k);
                }
                if (MK1(i, j + 1) <= k) {
                    MU(u, 2, i, j, k) = MU(u, 2, i, j, k) - sum[1] * MRH(2, i, j) + UBL(mo->ub, 2, i, j);
                    MU(u, 2, i, j, k) = mo->rel * MU(mo->u1, 2, i, j, k) + (1.0 - mo->rel) * MU(u, 2, i, j, k);
                    MU(mo->u1, 2, i, j, k) = MU(u, 2, i, j, k);
                }
            }
        }
    for (int j = 1; j <= MJ; j++)
        for (int k = MK1(1, j); k <= maxk; k++) MU(u, 1, 0, j, k) = MU(u, 1, MI, j, k);
    for (int j = 1; j <= MJ; j++)
        for (int i = 1; i <= MI; i++) {
            double tv = 0;
            for (int k = MK1(i, j); k <= maxk - 1; k++) {
                const double tv1 = (MU(u, 1, i, j, k) - MU(u, 1, i - 1, j, k)) * g->rdphi * g->rc[j];
                const double tv2 = (MU(u, 2, i, j, k) * g->cv[j] - MU(u, 2, i, j - 1, k) * g->cv[j - 1]) * g->rds[j];
                MU(u, 3, i, j, k) = tv - g->dz[k] * (tv1 + tv2);
                tv = MU(u, 3, i, j, k);
            }
        }
}

double og_mom_island(const og_mom *mo, const og_grid *g, int isl) { return mom_island(mo, g, mo->ub, isl, 1); }

double og_mom_island_of(const og_mom *mo, const og_grid *g, const double *ubv, int isl, int indj)
{
    return mom_island(mo, g, ubv, isl, indj);
}

/* ------------------------------------------------------------------------ */
/* EMBM surflux without ENTS (embm.f90:2548-3738)                            */
/* ------------------------------------------------------------------------ */
/* embm_lib.f90:36-157 (PARAMETERs, REAL*8) */
static const double E_RHOAIR = 1.25, E_RHO0 = 1.0E3, E_RHOAO = 1.25 / 1.0E3;
static const double E_CPA = 1004.0, E_HLV = 2.501E6, E_HLS = 2.501E6 + 3.34E5;  /* hls = hlv + hlf */
static const double E_CO20 = 278.0E-6, E_CH40 = 700.0E-9, E_N2O0 = 275.0E-9;
static const double E_ALPHACH4 = 0.036, E_ALPHAN2O = 0.12;
static const double E_B00 = 2.43414E2, E_B10 = -3.47968E1, E_B20 = 1.02790E1;
static const double E_B01 = 2.60065, E_B11 = -1.62064, E_B21 = 6.34856E-1;
static const double E_B02 = 4.40272E-3, E_B12 = -2.26092E-2, E_B22 = 1.12265E-2;
static const double E_B03 = -2.05237E-5, E_B13 = -9.67000E-5, E_B23 = 5.62925E-5;
static const double E_EMO = 0.94 * 5.67E-8, E_EMA = 0.85 * 5.67E-8;  /* 0.94, 0.85 sigma */
static const double E_TFREEZ = 0.0;
static const double E_CONST1 = 3.80E-3, E_CONST2 = 21.87, E_CONST3 = 265.5, E_CONST4 = 17.67, E_CONST5 = 243.5;
static const double E_CD = 0.0013;
static const double E_CONSIC = 2.166;
/* rhmin = 1/hmin, rhooi = rho0/rhoice, rrholf = 1/(rhoice hlf): hmin 0.01, rhoice 913, hlf 3.34E5 */
static const double E_RHMIN = 1.0 / 0.01, E_RHOOI = 1.0E3 / 913.0, E_RRHOLF = 1.0 / (913.0 * 3.34E5);
static const double E_M2MM = 1000.0, E_MM2M = 1.0 / 1000.0;
static const double E_HATMBL2 = 1800.;   /* hatmbl(2), embm.f90:1081 */
static const double E_SALN0 = 34.9, E_ZEROC = 273.15;
static const double E_CPSC = 3981.1;

#define F2(a, i, j) ((a)[(size_t)((i) - 1) + (size_t)p->maxi * (size_t)((j) - 1)])
#define SK1(i, j) (k1[(i) + (p->maxi + 2) * (j)])

/* ch4_func, embm.f90:3771-3777 */
static double sfx_ch4_func(double ch4, double n2o)
{
    return 0.47 * log(1.0 + 2.01E-5 * pow(ch4 * n2o, 0.75) + 5.31E-15 * ch4 * pow(ch4 * n2o, 1.52));
}

static double fmaxf_(double a, double b) { return b > a ? b : a; }  /* MAX(a, b) */
static double fminf_(double a, double b) { return b < a ? b : a; }  /* MIN(a, b) */

void og_sfx_usurf(const og_sfx_par *p, const og_grid *g, const double *sxu, const double *syu,
                  const double *sxv, const double *syv, double *usurf)
{
    const int maxi = p->maxi, maxj = p->maxj, maxk = p->maxk;
    const double dzz = g->dz[maxk] * g->dza[maxk - 1] / 2;           /* embm.f90:808 */
    const double den = RH0SC * DSC * USC * FSC;
    double *tau1 = malloc(sizeof(double) * maxi * maxj), *tau2 = malloc(sizeof(double) * maxi * maxj);
    /* 2756-2768 */
    for (int j = 1; j <= maxj; j++)
        for (int i = 1; i <= maxi; i++) {
            const double dztau1 = p->scf * F2(sxu, i, j) / den / dzz;
            const double dztav2 = p->scf * F2(syv, i, j) / den / dzz;
            F2(tau1, i, j) = dztau1 * dzz;
            F2(tau2, i, j) = dztav2 * dzz;
        }
    (void)syu;
    (void)sxv;
    /* 2789-2810 */
    for (int j = 1; j <= maxj; j++) {
        double tv3 = 0.0;
        for (int i = 1; i <= maxi; i++) {
            const double tv = i == 1 ? (F2(tau1, i, j) + F2(tau1, maxi, j)) / 2 : (F2(tau1, i, j) + F2(tau1, i - 1, j)) / 2;
            const double tv2 = j == 1 ? F2(tau2, i, j) / 2 : (F2(tau2, i, j) + F2(tau2, i, j - 1)) / 2;
            F2(usurf, i, j) = sqrt(sqrt(tv * tv + tv2 * tv2) * RH0SC * DSC * USC * FSC / (E_RHOAIR * E_CD * p->scf));
            tv3 = tv3 + F2(usurf, i, j);
        }
        if (p->polar_avg != 2)
            if (j <= 2 || j >= maxj - 1)
                for (int i = 1; i <= maxi; i++) F2(usurf, i, j) = tv3 / maxi;
    }
    free(tau1);
    free(tau2);
}

void og_surflux(const og_sfx_par *p, const og_grid *g, const int32_t *k1, og_sfx_state *st,
                const double *usurf, const double *otemp, const double *osaln, const double *pco2,
                double *out, int32_t *iters)
{
    const int maxi = p->maxi, maxj = p->maxj, maxk = p->maxk;
    const size_t nc = (size_t)maxi * maxj;
#define O(f, i, j) (out[(size_t)(f) * nc + (size_t)((i) - 1) + (size_t)maxi * (size_t)((j) - 1)])
    const double tol = 1.0E-10;      /* 2673 */
    const int itice = 21;            /* 2637 */
    const int64_t istot = st->istep + 1;   /* istep_ocn of this call (genie.f90:276) */
    const int imth = (int)((istot - 1) % p->nyear) + 1;
    /* greenhouse gases (2853-2926): option 1 or 4; the fields are uniform,
     * so one value stands for every cell */
    double co2 = st->co2, ch4 = st->ch4, n2o = st->n2o;
    if (!p->radfor_atchem) {
        co2 = (1.0 + p->rate_co2) * co2;
        ch4 = (1.0 + p->rate_ch4) * ch4;
        n2o = (1.0 + p->rate_n2o) * n2o;
        st->co2 = co2;
        st->ch4 = ch4;
        st->n2o = n2o;
    }
    double *runoff = calloc(nc, sizeof(double));
    double *pptn = calloc(nc, sizeof(double)), *evap = calloc(nc, sizeof(double));
    double *evapsic = calloc(nc, sizeof(double));
    /* 2931-2950 */
    memset(out, 0, sizeof(double) * OG_SF_N * nc);
    for (int j = 1; j <= maxj; j++)
        for (int i = 1; i <= maxi; i++) O(OG_SF_ALBEDO, i, j) = F2(p->albcl, i, j);
    /* 2953 */
    double meantemp = 0.0;
    for (int j = 1; j <= maxj; j++)
        for (int i = 1; i <= maxi; i++) meantemp = meantemp + F2(st->atemp, i, j);
    meantemp = meantemp / (double)(maxj * maxi);
    for (int i = 1; i <= maxi; i++)
        for (int j = 1; j <= maxj; j++) {
            double *ashum = &F2(st->ashum, i, j), *tice = &F2(st->tice, i, j);
            const double atemp = F2(st->atemp, i, j), sica = F2(st->sica, i, j), sich = F2(st->sich, i, j);
            const double us = F2(usurf, i, j), ca = F2(p->ca, i, j);
            if (iters) F2(iters, i, j) = 0;
            double cco2 = co2, cch4 = ch4, cn2o = n2o;
            if (p->radfor_atchem) {
                /* 2870-2887: dum_sfcatm(3) is the CO2; CH4 and N2O are not
                 * carried, so they read zero and take their references */
                const double pc = F2(pco2, i, j);
                cco2 = pc < E_CO20 / 1000.0 ? E_CO20 : pc;
                cch4 = E_CH40;
                cn2o = E_N2O0;
            }
            /* 2973-2991 */
            const double qsata = E_CONST1 * exp(E_CONST4 * atemp / (atemp + E_CONST5));
            F2(pptn, i, j) = fmaxf_(0.0, (*ashum - p->rmax * qsata) * E_RHOAO * E_HATMBL2 * p->rdtdim);
            /* 3000-3006 */
            *ashum = fminf_(*ashum, p->rmax * qsata);
            const double rq = *ashum / qsata;
            /* 3020-3024 */
            const double fxsw = p->solfor[(size_t)(j - 1) + (size_t)maxj * (size_t)(imth - 1)] * (1.0 - F2(p->albcl, i, j));
            /* 3047-3077 */
            const double tv0 = E_B00 + rq * (E_B10 + E_B20 * rq);
            const double tv1 = E_B01 + rq * (E_B11 + E_B21 * rq);
            const double tv2 = E_B02 + rq * (E_B12 + E_B22 * rq);
            const double tv3 = E_B03 + rq * (E_B13 + E_B23 * rq);
            const double ch4_term = E_ALPHACH4 * (sqrt(1.0E9 * cch4) - sqrt(1.0E9 * E_CH40)) -
                                    sfx_ch4_func(1.0E9 * cch4, 1.0E9 * E_N2O0) + sfx_ch4_func(1.0E9 * E_CH40, 1.0E9 * E_N2O0);
            const double n2o_term = E_ALPHAN2O * (sqrt(1.0E9 * cn2o) - sqrt(1.0E9 * E_N2O0)) -
                                    sfx_ch4_func(1.0E9 * E_CH40, 1.0E9 * cn2o) + sfx_ch4_func(1.0E9 * E_CH40, 1.0E9 * E_N2O0);
            const double fxplw = tv0 + atemp * (tv1 + atemp * (tv2 + atemp * tv3)) - p->delf2x * log(cco2 / E_CO20) -
                                 ch4_term - n2o_term + p->olr_adj * (meantemp - p->t_eqm) - p->olr_adj0;
            /* 3083 */
            const double fxlata = E_RHO0 * F2(pptn, i, j) * E_HLV;
            O(OG_SF_FXSW, i, j) = fxsw;
            O(OG_SF_PPTN, i, j) = F2(pptn, i, j);
            if (SK1(i, j) <= maxk) {
                const double ot = F2(otemp, i, j);
                /* 3088-3103 */
                double alw = atemp + E_ZEROC;
                alw = alw * alw;
                alw = alw * alw;
                alw = E_EMA * alw;
                const double salt = E_SALN0 + F2(osaln, i, j);
                const double tsfreez = salt * (-0.0575 + 0.0017 * sqrt(salt) - 0.0002 * salt);
                const double qb = p->rsictscsf * (tsfreez - ot);
                O(OG_SF_TSFREEZ, i, j) = tsfreez;
                double qbsic = qb;
                double albsic, fx0sica, dhsic, atm_latenti, atm_sensiblei, atm_netsoli, atm_netlongi;
                if (sica > 0.0) {
                    /* 3106-3227 */
                    albsic = fmaxf_(p->albsic_min, fminf_(p->albsic_max, 0.40 - 0.04 * atemp));
                    const double fxswsic = p->solfor[(size_t)(j - 1) + (size_t)maxj * (size_t)(imth - 1)] * (1.0 - albsic);
                    double cesic, chsic, cfxsensic, qsatsic;
                    for (int iter = 1; iter <= itice; iter++) {
                        if (iters) F2(iters, i, j) = iter;
                        const double ticold = *tice;
                        cesic = 1.0E-3 * (1.0022 - 0.0822 * (atemp - ticold) + 0.0266 * us);
                        cesic = fmaxf_(6.0E-5, fminf_(2.19E-3, cesic));
                        chsic = 0.94 * cesic;
                        cfxsensic = E_RHOAIR * chsic * E_CPA * us;
                        qsatsic = E_CONST1 * exp(E_CONST2 * ticold / (ticold + E_CONST3));
                        F2(evapsic, i, j) = fmaxf_(0.0, (qsatsic - *ashum) * E_RHOAO * cesic * us);
                        const double tk = ticold + E_ZEROC, tc3 = ticold + E_CONST3;
                        const double tieqn = sich * ((1 - ca) * fxswsic + alw - E_EMO * powi(tk, 4) -
                                                     cfxsensic * (ticold - atemp) - E_RHO0 * E_HLS * F2(evapsic, i, j)) +
                                             E_CONSIC * (tsfreez - ticold);
                        const double dtieq = sich * (-4.0 * E_EMO * powi(tk, 3) - cfxsensic -
                                                     E_HLS * E_RHOAIR * cesic * us * qsatsic * E_CONST2 * E_CONST3 /
                                                         powi(tc3, 2) * 0.5 * (1.0 + copysign(1.0, qsatsic - *ashum))) -
                                             E_CONSIC;
                        *tice = ticold - tieqn / dtieq;
                        if (fabs(*tice - ticold) < tol || (ticold > E_TFREEZ && tieqn > 0.0)) break;
                    }
                    /* label 10 */
                    *tice = fminf_(E_TFREEZ, *tice);
                    const double tk = *tice + E_ZEROC;
                    const double fxlwsic = E_EMO * powi(tk, 4) - alw;
                    cesic = 1.0E-3 * (1.0022 - 0.0822 * (atemp - *tice) + 0.0266 * us);
                    cesic = fmaxf_(6.0E-5, fminf_(2.19E-3, cesic));
                    chsic = 0.94 * cesic;
                    cfxsensic = E_RHOAIR * chsic * E_CPA * us;
                    const double fxsensic = cfxsensic * (*tice - atemp);
                    qsatsic = E_CONST1 * exp(E_CONST2 * *tice / (*tice + E_CONST3));
                    F2(evapsic, i, j) = fmaxf_(0.0, (qsatsic - *ashum) * E_RHOAO * cesic * us);
                    const double fx0sic = (1 - ca) * fxswsic - fxsensic - fxlwsic - E_RHO0 * E_HLS * F2(evapsic, i, j);
                    O(OG_SF_FX0SIC, i, j) = fx0sic;
                    fx0sica = ca * fxswsic + fxlata + fxsensic + fxlwsic - fxplw;
                    atm_latenti = +fxlata;
                    atm_sensiblei = +fxsensic;
                    atm_netsoli = +ca * fxswsic;
                    atm_netlongi = +fxlwsic - fxplw;
                    dhsic = E_RRHOLF * (qb - fx0sic) - E_RHOOI * F2(evapsic, i, j);
                    if (sich >= p->sich_max)
                        if (dhsic > 0.0) {
                            qbsic = (0.0 + E_RHOOI * F2(evapsic, i, j)) / E_RRHOLF + fx0sic;
                            dhsic = E_RRHOLF * (qbsic - fx0sic) - E_RHOOI * F2(evapsic, i, j);
                        }
                } else {
                    /* 3228-3238 */
                    albsic = 0.0;
                    fx0sica = 0.0;
                    dhsic = 0.0;
                    F2(evapsic, i, j) = 0.0;
                    *tice = 0.0;
                    atm_latenti = atm_sensiblei = atm_netsoli = atm_netlongi = 0.0;
                }
                /* 3239-3254 */
                const double tk = ot + E_ZEROC;
                const double fxlw = E_EMO * powi(tk, 4) - alw;
                double ce = 1.0E-3 * (1.0022 - 0.0822 * (atemp - ot) + 0.0266 * us);
                ce = fmaxf_(6.0E-5, fminf_(2.19E-3, ce));
                const double ch = 0.94 * ce;
                const double fxsen = E_RHOAIR * ch * E_CPA * us * (ot - atemp);
                const double qsato = E_CONST1 * exp(E_CONST4 * ot / (ot + E_CONST5));
                F2(evap, i, j) = fmaxf_(0.0, (qsato - *ashum) * E_RHOAO * ce * us);
                /* 3256-3277 */
                const double fx0oa = ca * fxsw + fxlata + fxsen + fxlw - fxplw;
                const double atm_latent = +fxlata, atm_sensible = +fxsen, atm_netsol = +ca * fxsw, atm_netlong = +fxlw - fxplw;
                O(OG_SF_FX0A, i, j) = (1 - sica) * fx0oa + sica * fx0sica;
                O(OG_SF_FXLHA, i, j) = (sica * atm_latenti) + ((1 - sica) * atm_latent);
                O(OG_SF_FXSHA, i, j) = (sica * atm_sensiblei) + ((1 - sica) * atm_sensible);
                O(OG_SF_FXSWA, i, j) = (sica * atm_netsoli) + ((1 - sica) * atm_netsol);
                O(OG_SF_FXLWA, i, j) = (sica * atm_netlongi) + ((1 - sica) * atm_netlong);
                /* 3279-3306 */
                const double fx0o = (1 - ca) * fxsw - fxsen - fxlw - E_RHO0 * E_HLV * F2(evap, i, j);
                O(OG_SF_FX0NETO, i, j) = sica * qbsic + (1 - sica) * fmaxf_(qb, fx0o);
                O(OG_SF_FXLHO, i, j) = (1 - sica) * (-(E_RHO0 * E_HLV * F2(evap, i, j)) + fmaxf_(0.0, qb - fx0o)) + sica * qbsic;
                O(OG_SF_FXSHO, i, j) = -((1 - sica) * fxsen);
                O(OG_SF_FXSWO, i, j) = (1 - sica) * (1 - ca) * fxsw;
                O(OG_SF_FXLWO, i, j) = -((1 - sica) * fxlw);
                const double dho = fmaxf_(0.0, E_RRHOLF * (qb - fx0o));
                O(OG_SF_DTHSIC, i, j) = sica * dhsic + (1 - sica) * dho;
                O(OG_SF_DTAREASIC, i, j) = fmaxf_(0.0, E_RHMIN * dho * (1 - sica));
                if (sich > 1.0E-12)
                    O(OG_SF_DTAREASIC, i, j) = O(OG_SF_DTAREASIC, i, j) + fminf_(0.0, 0.5 * sica * sica * dhsic / sich);
                /* 3324-3331 */
                O(OG_SF_ALBEDO, i, j) = sica * albsic + (1 - sica) * F2(p->albcl, i, j);
                O(OG_SF_ALBICE, i, j) = albsic;
                O(OG_SF_FX0O, i, j) = fx0o;
                O(OG_SF_FXSEN, i, j) = fxsen;
                O(OG_SF_FXLW, i, j) = fxlw;
            } else {
                /* 3332-3360 */
                O(OG_SF_FX0A, i, j) = fxsw + fxlata - fxplw;
                O(OG_SF_FXLHA, i, j) = +fxlata;
                O(OG_SF_FXSHA, i, j) = +0.0;
                O(OG_SF_FXSWA, i, j) = +fxsw;
                O(OG_SF_FXLWA, i, j) = -fxplw;
                const int ir = F2(p->iroff, i, j), jr = F2(p->jroff, i, j);
                if (p->igrid != 0)
                    F2(runoff, ir, jr) = F2(runoff, ir, jr) + F2(pptn, i, j) * g->ds[j] * g->rds[jr];
                else
                    F2(runoff, ir, jr) = F2(runoff, ir, jr) + F2(pptn, i, j);
                O(OG_SF_RUNOFF_LAND, i, j) = F2(pptn, i, j);
            }
        }
    /* 3640-3679 */
    for (int j = 1; j <= maxj; j++)
        for (int i = 1; i <= maxi; i++) {
            const double sica = F2(st->sica, i, j);
            if (SK1(i, j) <= maxk) {
                O(OG_SF_PPTN_OCN, i, j) = F2(pptn, i, j);
                O(OG_SF_RUNOFF_OCN, i, j) = F2(runoff, i, j) + 0.0;
                O(OG_SF_EVAP_ATM, i, j) = F2(evap, i, j) * (1 - sica) + F2(evapsic, i, j) * sica;
                O(OG_SF_PPTN_OCN, i, j) = O(OG_SF_PPTN_OCN, i, j) + F2(p->pmeadj, i, j);
            } else {
                O(OG_SF_PPTN_OCN, i, j) = 0.0;
                O(OG_SF_RUNOFF_OCN, i, j) = 0.0;
                O(OG_SF_EVAP_ATM, i, j) = 0.0;
            }
            O(OG_SF_PPTN_ATM, i, j) = F2(pptn, i, j);
            O(OG_SF_EVAP_OCN, i, j) = -O(OG_SF_EVAP_ATM, i, j);
            O(OG_SF_PPTN_OCN, i, j) = O(OG_SF_PPTN_OCN, i, j) * E_M2MM;
            O(OG_SF_EVAP_OCN, i, j) = O(OG_SF_EVAP_OCN, i, j) * E_M2MM;
            O(OG_SF_RUNOFF_OCN, i, j) = O(OG_SF_RUNOFF_OCN, i, j) * E_M2MM;
            O(OG_SF_RUNOFF_LAND, i, j) = O(OG_SF_RUNOFF_LAND, i, j) * E_M2MM;
            O(OG_SF_PPTN_ATM, i, j) = O(OG_SF_PPTN_ATM, i, j) * E_M2MM;
            O(OG_SF_EVAP_ATM, i, j) = O(OG_SF_EVAP_ATM, i, j) * E_M2MM;
            O(OG_SF_EVAP, i, j) = F2(evap, i, j);
        }
    st->istep = istot;
    free(runoff);
    free(pptn);
    free(evap);
    free(evapsic);
#undef O
}

void og_surface_slot(const og_sfx_par *p, const double *out, const double *fx0sic, const double *fwsic,
                     int maxl, double *sflux)
{
    const size_t nc = (size_t)p->maxi * p->maxj;
    /* goldstein_lib.f90:63 and goldstein.f90:904 */
    const double rfluxsc = RSC / (DSC * USC * RH0SC * E_CPSC);
    const double rpmesco = RSC * E_SALN0 / (DSC * USC);
    for (size_t c = 0; c < nc; c++) {
        /* 149-170; fw_hosing and fw_anom are zero without hosing or an anomaly file */
        const double fx0neto = out[OG_SF_FXSWO * nc + c] + out[OG_SF_FXSHO * nc + c] + out[OG_SF_FXLWO * nc + c] +
                               out[OG_SF_FXLHO * nc + c] + (fx0sic ? fx0sic[c] : 0.0);
        double fwfxneto = out[OG_SF_PPTN_OCN * nc + c] + out[OG_SF_EVAP_OCN * nc + c] + out[OG_SF_RUNOFF_OCN * nc + c] +
                          (fwsic ? fwsic[c] : 0.0) + 0.0 + 0.0;
        fwfxneto = fwfxneto * E_MM2M;
        sflux[0 + (size_t)maxl * c] = -fx0neto * rfluxsc;
        sflux[1 + (size_t)maxl * c] = fwfxneto * rpmesco;
    }
}
#undef F2
#undef SK1
