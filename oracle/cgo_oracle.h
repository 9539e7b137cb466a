/*
 * cgo_oracle.h -- CPU restatement of the cGENIE ocean tracer timestep.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker for the HIP engine
 * in cgenie_amd/.  Only tests/, __graft_entry__.smoke() and the cpu_baseline
 * leg of bench.py may load it.  The product path never calls it.
 *
 * Every routine restates one reference routine with the reference's loop
 * order, carried fluxes and operand order (file:line cited at each function
 * in cgo_oracle.c).  All arrays use the reference's own Fortran layouts
 * (column-major, lower bounds as in the reference ALLOCATE statements):
 *
 *   go_ts  (maxl, maxi, maxj, maxk)            goldstein.f90:31, 443
 *   k1     (0:maxi+1, 0:maxj+1)  int32         goldstein_lib.f90:9, goldstein.f90:655
 *   u      (3, 0:maxi, 0:maxj, maxk)           goldstein.f90:684
 *   sflux  (maxl, maxi, maxj)  = ts(:,1:maxi,1:maxj,maxk+1)   goldstein.f90:167-170
 *   rho    (0:maxi+1, 0:maxj+1, 0:maxk)        goldstein.f90:704
 */
#ifndef CGO_ORACLE_H
#define CGO_ORACLE_H
#include <stdint.h>
#ifdef __cplusplus
extern "C" {
#endif

#define OG_MAXJ 256
#define OG_MAXK 64

typedef struct og_grid {
    int maxi, maxj, maxk;
    double dphi, rdphi;
    /* horizontal metrics, index 0..maxj+1 (Fortran index == C index) */
    double s[OG_MAXJ + 2], sv[OG_MAXJ + 2], c[OG_MAXJ + 2], cv[OG_MAXJ + 2];
    double ds[OG_MAXJ + 2], rds[OG_MAXJ + 2], rc[OG_MAXJ + 2], rc2[OG_MAXJ + 2];
    double dsv[OG_MAXJ + 2], rdsv[OG_MAXJ + 2], rcv[OG_MAXJ + 2], cv2[OG_MAXJ + 2];
    double rds2[OG_MAXJ + 2];
    /* vertical metrics, index 0..maxk+1 */
    double dz[OG_MAXK + 2], dza[OG_MAXK + 2], rdz[OG_MAXK + 2], rdza[OG_MAXK + 2];
    double zw[OG_MAXK + 2], zro[OG_MAXK + 2], dt[OG_MAXK + 2], ssmax[OG_MAXK + 2];
    double ec[5];
    double diff[2];   /* non-dimensional horizontal / vertical diffusivity */
    double tsc, dtyr; /* time scale (s); ocean step in years */
} og_grid;

/* goldstein.f90:906-1075, 1406-1407, 2057-2070 */
void og_grid_init(og_grid *g, int maxi, int maxj, int maxk, int nyear,
                  double yearlen, double kappa_h, double kappa_v,
                  double ssmaxsurf, double ssmaxdeep);

/* One step_goldstein tracer step (goldstein.f90:149-187, 235, 443-444):
 * builds halo'd ts/ts1 from go_ts, runs tstepo (flux, co, halos, ts1=ts) and
 * writes the interior back.  rho is in/out (its wet entries are rewritten),
 * cost is accumulated, limps/dmax are the tstepo_flux diagnostics. */
void og_goldstein_step(const og_grid *g, int maxl, const int32_t *k1,
                       double *go_ts, const double *u, const double *sflux,
                       double *rho, double *cost, int32_t *mldpk, int diso,
                       int32_t *limps, double *dmax);

/* Per-tracer volume-weighted inventory over wet cells, the diag2 weight
 * dphi*ds(j)*dz(k) (goldstein_diag.f90:39-41).  inv has maxl entries. */
void og_inventory(const og_grid *g, int maxl, const int32_t *k1,
                  const double *go_ts, double *inv);

/* eos / eosd, goldstein.f90:3048-3082 (ieos = 0 form) */
double og_eos(const double *ec, double t, double s);
void og_eosd(const double *ec, double t1, double t2, double s1, double s2,
             double rdz, double *dzrho, double *tec);

/* ---------------- carbonate chemistry (gem_carbchem.f90) ---------------- */
enum { OG_ICC_K = 0, OG_ICC_K1, OG_ICC_K2, OG_ICC_KB, OG_ICC_KW, OG_ICC_KSI,
       OG_ICC_KHF, OG_ICC_KHSO4, OG_ICC_KP1, OG_ICC_KP2, OG_ICC_KP3,
       OG_ICC_KCAL, OG_ICC_KARG, OG_ICC_QCO2, OG_ICC_QO2, OG_ICC_KH2S,
       OG_ICC_KNH4, OG_N_CARBCONST };
enum { OG_IC_H = 0, OG_IC_FUG_CO2, OG_IC_CONC_CO2, OG_IC_CONC_CO3,
       OG_IC_CONC_HCO3, OG_IC_OHM_CAL, OG_IC_OHM_ARG, OG_IC_DCO3_CAL,
       OG_IC_DCO3_ARG, OG_IC_RF0, OG_N_CARB };
#define OG_N_CARBALK 13

/* sub_calc_carbconst, Mehrbach default set (gem_carbchem.f90:59-268) */
void og_calc_carbconst(double D, double T, double S, double *cc);
/* salinity-derived totals (gem_carbchem.f90:834-890) */
double og_calc_Ca(double S);
double og_calc_Btot(double S);
double og_calc_SO4tot(double S);
double og_calc_Ftot(double S);
/* sub_calc_carb fixed point (gem_carbchem.f90:326-546).  carb[OG_IC_H] is the
 * warm start on entry.  Returns the iteration count (>0) or a negative code:
 * -1 negative [H] estimate, -2 iteration limit.  On failure carb is left as on
 * entry (the reference's random re-seed path, 468-476/527-536, is not taken). */
int og_calc_carb(double DIC, double ALK, double Ca, double PO4, double SiO2,
                 double Btot, double SO4tot, double Ftot, double H2S,
                 double NH4, const double *cc, double *carb, double *carbalk,
                 double tol, int maxit);
/* sub_calc_carb_RF0 (gem_carbchem.f90:553-669); writes carb[OG_IC_RF0] */
int og_calc_carb_RF0(double DIC, double ALK, double PO4, double SiO2,
                     double Btot, double SO4tot, double Ftot, double H2S,
                     double NH4, const double *cc, double *carb, double tol,
                     int maxit);

/* ---------------- BIOGEM (biogem.f90, biogem_box.f90) ---------------- */
/* compact sediment-tracer slots, in the reference's is order
 * (is_POC=3, is_POP=8, is_CaCO3=14, is_POC_frac2=33, is_CaCO3_frac2=34) */
enum { OG_S_POC = 0, OG_S_POP, OG_S_CACO3, OG_S_POCFR2, OG_S_CACO3FR2, OG_NSED };

typedef struct og_bgc_params {
    /* compact ocean-tracer slot of each selected species, -1 if absent;
     * slot 0 = T, slot 1 = S as in the GOLDSTEIN ts array */
    int l_DIC, l_ALK, l_PO4, l_O2, l_DOMC, l_DOMP;
    int caco3;            /* CaCO3 particulate selected (needs DIC and ALK) */
    double dtyr;          /* BIOGEM step (yr) */
    double k0_PO4, c0_PO4;                /* biogem-defaults.nml:235-237 */
    double DOMfrac, DOMlifetime;          /* 0.66, 0.5 yr */
    double red_POP_POC, red_POP_PO2, red_PON_ALK, red_POP_PON;
    double red_POC_CaCO3, red_POC_CaCO3_pP;
    double POC_frac2, POC_dfrac2, POC_c0frac2, POC_eL1, POC_eL2;
    double CaCO3_frac2, CaCO3_eL1, CaCO3_eL2;
    double sinkrate;      /* m/yr */
    double remin_k_O2, remin_c0_O2;
    double carb_tol;      /* par_carbchem_pH_tolerance */
    int carb_maxit;
    double dsc;           /* goldstein_dsc (m) */
    /* air-sea exchange of CO2 and O2 with a well-mixed atmosphere */
    int airsea;           /* on/off */
    double gastransfer_a; /* par_gastransfer_a, biogem-defaults.nml:221 */
    double atm_pCO2, atm_pO2; /* initial atmospheric partial pressures (atm) */
} og_bgc_params;

/* Reference defaults (biogem-defaults.nml:233-303, gem-defaults.nml:196-199). */
void og_bgc_default_params(og_bgc_params *p, double dtyr);

typedef struct og_bgc_state {
    int maxi, maxj, maxk, maxl;
    /* ocn (maxl, maxi, maxj, maxk): T in K, S absolute, BGC tracers absolute */
    double *ocn;
    double *part;     /* (OG_NSED, maxi, maxj, maxk) bio_part */
    double *settle;   /* (OG_NSED, maxi, maxj, maxk) bio_settle */
    double *carb;     /* (OG_N_CARB, maxi, maxj) surface carb */
    double *carbconst;/* (OG_N_CARBCONST, maxi, maxj) surface carbconst */
    double *M, *rM, *V; /* (maxi, maxj, maxk) phys_ocn ipo_M/rM/V */
    double Dbot[OG_MAXK + 2], Dmid[OG_MAXK + 2], dD[OG_MAXK + 2]; /* by k */
    double *light;    /* (maxi, maxj) solfor/solar_constant */
    double *mld;      /* (maxi, maxj) mixed-layer depth, m */
    int32_t carb_fail; /* count of cells whose carbonate solve failed */
    int32_t carb_iter; /* iterations of the last surface solve sweep (sum) */
    /* with air-sea exchange: (2, maxi, maxj) CO2 then O2 of each column */
    double *wspeed;   /* (maxi, maxj) wind speed (m/s), par_phys_windspeed */
    double *atm;      /* atmospheric partial pressure (atm), dum_sfcatm1 */
    double *fatm;     /* this step's ocean -> atmosphere flux (mol yr-1) */
    double *sumatm;   /* the flux integrated since the last ATCHEM step (mol m-2) */
} og_bgc_state;

/* sub_init_phys_ocn geometry (biogem_data.f90:1098-1140) and ocn from go_ts
 * (ocn = ts*S/Smean for BGC and dye slots, T+273.15, S+34.9);
 * carb H start 10**(-7.8) and a surface solve (biogem_data.f90:2336-2432). */
void og_bgc_init(og_bgc_state *st, const og_grid *g, const og_bgc_params *p,
                 int maxl, const int32_t *k1, const double *go_ts,
                 const double *light, const double *mld);
void og_bgc_free(og_bgc_state *st);

/* fun_calc_rho, gem_util.f90:670-688 (T in K) */
double og_calc_rho(double T, double S);
/* One surface cell's air-sea exchange as step_biogem forms it
 * (biogem.f90:1044-1046, 1605; biogem_box.f90: sub_calc_solconst 27-40,
 * sub_calc_pv 81-117, fun_calc_ocnatm_flux 123-299; gem_carbchem.f90:
 * fun_calc_solconst 1367-1426): out[0..1] solubility of CO2 and O2,
 * out[2..3] piston velocity, out[4..5] ocean -> atmosphere flux (mol yr-1). */
void og_airsea_cell(double T, double S, double rho, double dD, double A, double wspeed, double DIC, double O2,
                    double CO2, double RF0, double pCO2, double pO2, double dtyr, double gas_a, double *out);
/* wind speed (maxi, maxj) for the air-sea exchange */
void og_bgc_set_windspeed(og_bgc_state *st, const double *wspeed);
/* the ocean -> atmosphere coupling and ATCHEM step after a BIOGEM step
 * (cpl_flux_ocnatm, atchem.f90:306-320; step_atchem, atchem.f90:63-158):
 * well-mixed atmosphere over the ocean grid at 273.15 K */
void og_atchem_step(og_bgc_state *st, const og_grid *g, const og_bgc_params *p);

/* ---------------- GOLDSTEIN momentum: the velocity solve ----------------
 * Barotropic streamfunction with islands and baroclinic velocities
 * (goldstein.f90: invert 3138-3204, jbar 3217-3315, ubarsolv 3500-3565,
 * velc 3568-3679, wind 3685-3714, matinv_gold/matmult 3452-3492, drgset
 * 2845-2885, set-up 1359-1530, 1788-1814, the step 198-233;
 * goldstein_lib.f90: island 186-241).  Island data as the reference reads it:
 * labels(maxi, 0:maxj) of the .psiles file (landmass numbers at psi points,
 * 1 = the first landmass) and, per island, the path points (lpi, ipi, jpi)
 * of the .paths file. */
typedef struct og_mom {
    int maxi, maxj, maxk, isles, mpi;
    double rel;
    double *h, *rh;          /* (3, 0:maxi+1, 0:maxj+1) */
    int32_t *ku;             /* (2, maxi, maxj) */
    int32_t *mk;             /* (0:maxi+1, 0:maxj+1) */
    uint8_t *getj;           /* (maxi, maxj) */
    double *drag;            /* (2, 0:maxi+1, 0:maxj+1) */
    double *rtv, *rtv3;      /* (maxi, maxj) */
    double *gap, *ratm;      /* (N, 2n+3) and (N, n+1), N = maxi (maxj+1), n = maxi */
    double *gb, *gbold;      /* (N) */
    double *bp;              /* (0:maxi+1, 0:maxj+1, 0:maxk) */
    double *sbp;             /* (0:maxi+1, 0:maxj+1) */
    double *ub;              /* (2, 0:maxi+1, 0:maxj) */
    double *psi;             /* (0:maxi, 0:maxj) */
    double *ubisl, *psisl;   /* (2, 0:maxi+1, 0:maxj, isles), (0:maxi, 0:maxj, isles) */
    double *erisl, *psibc;   /* (isles, isles+1) column-major, (isles) */
    int32_t *npi, *lpisl, *ipisl, *jpisl; /* (isles), (mpi, isles) */
    double *tau, *dztau, *dztav;          /* (2, maxi, maxj) */
    double *u1;              /* (3, 0:maxi, 0:maxj, maxk), the relaxation memory */
} og_mom;

/* Set-up of initialise_goldstein for the momentum: depths, drag (adrag in
 * days, goldstein-defaults.nml:21), the banded factorisation of the psi
 * equation and the island basis.  npi[isles] path lengths, then the paths
 * concatenated as (lpi, ipi, jpi) triples.  Returns 0, or -1 for bad island
 * data (a dry or out-of-range path point, as the reference's checks). */
int og_mom_init(og_mom *mo, const og_grid *g, const int32_t *k1, const int32_t *labels, int isles,
                const int32_t *npi, const int32_t *paths, double adrag, double rel);
void og_mom_free(og_mom *mo);
/* Wind stress (N m-2): stress(4, maxi, maxj) = x at u points, y at u points,
 * x at v points, y at v points (stressxu/yu/xv/yv_ocn); scf the namelist's
 * wind scale (goldstein.f90:102-107) */
void og_mom_set_wind(og_mom *mo, const og_grid *g, const double *stress, double scf);
/* One momentum step of step_goldstein (goldstein.f90:198-233): wind, jbar,
 * ubarsolv, the island correction, velc.  rho (0:maxi+1, 0:maxj+1, 0:maxk)
 * as tstepo leaves it; u (3, 0:maxi, 0:maxj, maxk) in/out. */
void og_mom_step(og_mom *mo, const og_grid *g, const int32_t *k1, const double *rho, double *u);
/* island(ub, ., isl, 1) with the current barotropic velocity: the
 * circulation round island isl that the correction drives to zero */
double og_mom_island(const og_mom *mo, const og_grid *g, int isl);
/* the same path integral of a given ub (2, 0:maxi+1, 0:maxj); indj = 0 leaves
 * out the wind and the bottom-pressure torque */
double og_mom_island_of(const og_mom *mo, const og_grid *g, const double *ubv, int isl, int indj);
/* rho of every wet cell and the periodic halo from go_ts (eos at
 * initialisation, goldstein.f90:1430-1444, 1760-1765) */
void og_rho_from_ts(const og_grid *g, int maxl, const int32_t *k1, const double *go_ts, double *rho);

/* ---------------------------------------------------------------------
 * EMBM surface fluxes, surflux (embm.f90:2548-3738) without ENTS
 * (flag_ents = .false.): no orography, ice-sheet or time-series forcing, the
 * greenhouse gases compounding at their namelist rates (option 4) or CO2 read
 * from the atmosphere composition (option 1, atchem_radfor).  Set-up tables
 * as initialise_embm leaves them; fields (maxi, maxj), column-major.
 * --------------------------------------------------------------------- */
typedef struct og_sfx_par {
    int maxi, maxj, maxk, nyear;
    int igrid;          /* 0: runoff added as is; else weighted by ds(j)/ds(jroff) */
    int polar_avg;      /* par_wind_polar_avg: 2 leaves the polar rows' usurf alone */
    int radfor_atchem;  /* atchem_radfor = 'y': CO2 from the atmosphere composition */
    double scf, rmax, delf2x, rdtdim, rsictscsf, olr_adj, olr_adj0, t_eqm;
    double albsic_min, albsic_max, sich_max, rate_co2, rate_ch4, rate_n2o;
    const double *solfor;            /* solfor(maxj, nyear) */
    const double *albcl, *ca, *pmeadj;
    const int32_t *iroff, *jroff;    /* runoff destination of every cell (1-based) */
} og_sfx_par;

/* outputs of one call, each a (maxi, maxj) field, in this order */
enum {
    OG_SF_FXLHO = 0, OG_SF_FXSHO, OG_SF_FXSWO, OG_SF_FXLWO, OG_SF_EVAP_OCN, OG_SF_PPTN_OCN,
    OG_SF_RUNOFF_OCN, OG_SF_RUNOFF_LAND, OG_SF_FXLHA, OG_SF_FXSHA, OG_SF_FXSWA, OG_SF_FXLWA,
    OG_SF_EVAP_ATM, OG_SF_PPTN_ATM, OG_SF_DTHSIC, OG_SF_DTAREASIC, OG_SF_ALBEDO, OG_SF_ALBICE,
    OG_SF_FXSW, OG_SF_FX0A, OG_SF_FX0O, OG_SF_FXSEN, OG_SF_FXLW, OG_SF_EVAP, OG_SF_PPTN,
    /* EMBM's own arrays of the call over the ocean: fx0sic (0 without ice), fx0neto, tsfreez */
    OG_SF_FX0SIC, OG_SF_FX0NETO, OG_SF_TSFREEZ, OG_SF_N
};

/* The state surflux carries between calls: air temperature and specific
 * humidity (tq, clamped by each call), the sea-ice height, cover and
 * surface temperature (tice is the Newton iteration's start), the
 * greenhouse concentrations and the call count istep_ocn. */
typedef struct og_sfx_state {
    double *atemp, *ashum, *sich, *sica, *tice;   /* (maxi, maxj) */
    double co2, ch4, n2o;
    int64_t istep;
} og_sfx_state;

/* usurf (maxi, maxj) from the four stress fields (embm.f90:2756-2810) */
void og_sfx_usurf(const og_sfx_par *p, const og_grid *g, const double *sxu, const double *syu,
                  const double *sxv, const double *syv, double *usurf);
/* One surflux call.  otemp, osaln: the ocean's top-level T and S
 * (tstar_ocn, sstar_ocn); pco2: the atmosphere's pCO2 per cell (atm) when
 * radfor_atchem, else ignored; out: OG_SF_N fields; iters: the sea-ice
 * Newton iterations of every cell (or NULL). */
void og_surflux(const og_sfx_par *p, const og_grid *g, const int32_t *k1, og_sfx_state *st,
                const double *usurf, const double *otemp, const double *osaln, const double *pco2,
                double *out, int32_t *iters);
/* The surface slot ts(1:2, 1:maxi, 1:maxj, maxk+1) that step_goldstein
 * forms from the flux components (goldstein.f90:149-170); fx0sic, fwsic the
 * sea-ice model's heat and freshwater into the ocean (NULL = zero);
 * sflux (maxl, maxi, maxj), only its first two tracers written. */
void og_surface_slot(const og_sfx_par *p, const double *out, const double *fx0sic, const double *fwsic,
                     int maxl, double *sflux);

/* step_biogem restricted to the selected species (biogem.f90:596-611,
 * 873-969, 985-1105, 1693, 1811-1844).  Writes vdocn (maxl,maxi,maxj,maxk). */
void og_biogem_step(og_bgc_state *st, const og_grid *g, const og_bgc_params *p,
                    const int32_t *k1, double *vdocn);

/* biogem_tracercoupling (biogem.f90:1885-2077); go_ts in/out */
void og_tracercoupling(og_bgc_state *st, const og_grid *g, const int32_t *k1,
                       double *go_ts, const double *vdocn);

#ifdef __cplusplus
}
#endif
#endif
